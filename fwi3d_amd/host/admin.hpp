// admin.hpp -- parameter file and raw-file helpers of the FWI3D drivers.
//
// Same interface and semantics as the reference's src/admin/admin.hpp:21-30
// (catch_parameter's prefix match, '#'/' ' line skipping, value taken from
// two characters after '=' up to '#', spaces removed), with two guards the
// reference lacks: empty lines are skipped instead of dereferencing
// line.front() (admin.cpp:68), and import_binary_float throws when a file
// holds fewer values than requested instead of silently under-reading
// (admin.cpp:20).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

void import_binary_float(std::string path, float* array, long n);
void export_binary_float(std::string path, const float* array, long n);

void import_text_file(std::string path, std::vector<std::string>& elements);

std::string catch_parameter(std::string target, std::string file);

// catch_parameter with a default for keys this build adds (absent from the
// reference's parameters.txt).
std::string catch_parameter_or(std::string target, std::string file, std::string fallback);

std::vector<std::string> split(std::string s, char delimiter);
