// geometry.hpp -- acquisition geometry (src/geometry/geometry.hpp:6-24):
// SPS / RPS text files of "x,y,z" lines in metres.
#pragma once

#include <string>
#include <vector>

class Geometry {
public:
    int nsrc = 0;
    int nrec = 0;

    std::vector<float> xsrc, ysrc, zsrc;
    std::vector<float> xrec, yrec, zrec;

    std::string parameters;

    void set_parameters();  // geometry.cpp:3-47
};
