// admin.cpp -- see admin.hpp.
#include "admin.hpp"

#include <algorithm>
#include <fstream>
#include <iostream>
#include <sstream>

// src/admin/admin.cpp:13-24 (plus a short-read check)
void import_binary_float(std::string path, float* array, long n) {
    std::ifstream file(path, std::ios::in | std::ios::binary);
    if (!file.is_open()) throw std::invalid_argument("Error: " + path + " could not be opened!");
    file.read((char*)array, n * (long)sizeof(float));
    if (file.gcount() != n * (long)sizeof(float))
        throw std::invalid_argument("Error: " + path + " holds fewer than " + std::to_string(n) + " floats!");
}

// src/admin/admin.cpp:26-37
void export_binary_float(std::string path, const float* array, long n) {
    std::ofstream file(path, std::ios::out | std::ios::binary);
    if (!file.is_open()) throw std::invalid_argument("Error: " + path + " could not be opened!");
    file.write((const char*)array, n * (long)sizeof(float));
    std::cout << "\nBinary file " << path << " was successfully written." << std::endl;
}

// src/admin/admin.cpp:39-51
void import_text_file(std::string path, std::vector<std::string>& elements) {
    std::ifstream file(path, std::ios::in);
    if (!file.is_open()) throw std::invalid_argument("Error: " + path + " could not be opened!");
    std::string line;
    while (std::getline(file, line))
        if (!line.empty() && line[0] != '#') elements.push_back(line);
}

// src/admin/admin.cpp:53-88: first line that *starts with* target wins.
std::string catch_parameter(std::string target, std::string file) {
    std::ifstream parameters(file);
    if (!parameters.is_open()) throw std::invalid_argument("Error: " + file + " could not be opened!");
    std::string line, variable;
    while (std::getline(parameters, line)) {
        if (line.empty() || line.front() == '#' || line.front() == ' ') continue;
        if (line.find(target) == 0) {
            const size_t eq = line.find('=');
            if (eq != std::string::npos)
                for (size_t i = eq + 2; i < line.size(); i++) {
                    if (line[i] == '#') break;
                    variable += line[i];
                }
            break;
        }
    }
    variable.erase(std::remove(variable.begin(), variable.end(), ' '), variable.end());
    variable.erase(std::remove(variable.begin(), variable.end(), '\r'), variable.end());
    return variable;
}

std::string catch_parameter_or(std::string target, std::string file, std::string fallback) {
    std::string v = catch_parameter(target, file);
    return v.empty() ? fallback : v;
}

// src/admin/admin.cpp:90-100
std::vector<std::string> split(std::string s, char delimiter) {
    std::string token;
    std::vector<std::string> tokens;
    std::istringstream tokenStream(s);
    while (std::getline(tokenStream, token, delimiter)) tokens.push_back(token);
    return tokens;
}
