// geometry.cpp -- Geometry::set_parameters (src/geometry/geometry.cpp:3-47).
#include "geometry.hpp"

#include <stdexcept>

#include "admin.hpp"

namespace {

void read_points(const std::string& path, std::vector<float>& x, std::vector<float>& y, std::vector<float>& z) {
    std::vector<std::string> lines;
    import_text_file(path, lines);
    for (const std::string& l : lines) {
        if (l.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::vector<std::string> aux = split(l, ',');
        if (aux.size() < 3) throw std::invalid_argument("Error: bad geometry line in " + path + ": " + l);
        x.push_back(std::stof(aux[0]));
        y.push_back(std::stof(aux[1]));
        z.push_back(std::stof(aux[2]));
    }
}

}  // namespace

void Geometry::set_parameters() {
    read_points(catch_parameter("SPS", parameters), xsrc, ysrc, zsrc);
    read_points(catch_parameter("RPS", parameters), xrec, yrec, zrec);
    nsrc = (int)xsrc.size();
    nrec = (int)xrec.size();
}
