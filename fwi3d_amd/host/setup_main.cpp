// setup_main -- acquisition files and initial FWI model of the reference
// workflow (run/generate_setup.py), without Python.
// Usage: setup_main parameters.txt
//
// Reads SPS, RPS, model_file, x/y/z_samples, model_spacing and depth_mask;
// writes the SPS/RPS files and model_file with every "true" replaced by
// "init" (1 / G_sigma(1 / v_true), top rows above depth_mask kept).
// Optional keys (defaults: the reference script's constants):
//   setup_shots = nsx,nsy (6,6)          setup_shot_range = x0,x1,y0,y1 (1000,4000,1000,4000)
//   setup_shot_depth = z (100)           setup_receivers = nrx,nry (51,51)
//   setup_receiver_range = x0,x1,y0,y1 (0,5000,0,5000)
//   setup_receiver_depth = z (0)         setup_smoothing = sigma in cells (3)
#include <iostream>
#include <vector>

#include "admin.hpp"
#include "dataset.hpp"

namespace {

std::vector<double> numbers(const std::string& key, const std::string& file, const std::string& fallback,
                            size_t count) {
    std::vector<double> v;
    for (const std::string& t : split(catch_parameter_or(key, file, fallback), ',')) v.push_back(std::stod(t));
    if (v.size() != count) throw std::invalid_argument(key + ": expected " + std::to_string(count) + " values");
    return v;
}

std::string replace_all(std::string s, const std::string& a, const std::string& b) {
    for (size_t p = s.find(a); p != std::string::npos; p = s.find(a, p + b.size())) s.replace(p, a.size(), b);
    return s;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt\n";
        return 1;
    }
    try {
        const std::string par = argv[1];
        Acquisition a;
        const auto ns = numbers("setup_shots", par, "6,6", 2);
        const auto sr = numbers("setup_shot_range", par, "1000,4000,1000,4000", 4);
        const auto nr = numbers("setup_receivers", par, "51,51", 2);
        const auto rr = numbers("setup_receiver_range", par, "0,5000,0,5000", 4);
        a.nsx = (int)ns[0];
        a.nsy = (int)ns[1];
        a.sx0 = sr[0], a.sx1 = sr[1], a.sy0 = sr[2], a.sy1 = sr[3];
        a.sz = std::stod(catch_parameter_or("setup_shot_depth", par, "100"));
        a.nrx = (int)nr[0];
        a.nry = (int)nr[1];
        a.rx0 = rr[0], a.rx1 = rr[1], a.ry0 = rr[2], a.ry1 = rr[3];
        a.rz = std::stod(catch_parameter_or("setup_receiver_depth", par, "0"));
        const std::string sps = catch_parameter("SPS", par), rps = catch_parameter("RPS", par);
        write_acquisition(sps, rps, a);

        const std::string model_file = catch_parameter("model_file", par);
        const int nx = std::stoi(catch_parameter("x_samples", par));
        const int ny = std::stoi(catch_parameter("y_samples", par));
        const int nz = std::stoi(catch_parameter("z_samples", par));
        const float dh = std::stof(catch_parameter("model_spacing", par));
        const float zmask = std::stof(catch_parameter("depth_mask", par));
        const double sigma = std::stod(catch_parameter_or("setup_smoothing", par, "3"));
        const long n = (long)nx * ny * nz;
        std::vector<float> m_true(n), m_init(n);
        import_binary_float(model_file, m_true.data(), n);
        initial_model(m_true.data(), m_init.data(), nz, nx, ny, dh, zmask, sigma);
        const std::string init_file = replace_all(model_file, "true", "init");
        export_binary_float(init_file, m_init.data(), n);
        std::cout << "SPS: " << sps << " (" << a.nsx * a.nsy << " shots)\nRPS: " << rps << " ("
                  << a.nrx * a.nry << " receivers)\ninitial model: " << init_file << std::endl;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    return 0;
}
