// preprocess_main -- RTM input data of the reference workflow
// (run/generate_input_data.py), without Python.
// Usage: preprocess_main parameters.txt
//
// For every shot of SPS, reads mod_output_folder/seismogram_nt{nt}_nr{nr}_
// {dt in us}us_shot_{s}.bin, normalises it by its largest |sample|, tapers
// it before the direct wave and writes input_RTM_<same name> beside it (the
// files migration_main reads through mig_input_folder).
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include "admin.hpp"
#include "dataset.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt\n";
        return 1;
    }
    try {
        const std::string par = argv[1];
        const int nt = std::stoi(catch_parameter("time_samples", par));
        const std::string dts = catch_parameter("time_spacing", par);
        const float dt = std::stof(dts);
        const std::string folder = catch_parameter("mod_output_folder", par);
        std::vector<float> sx, sy, sz, rx, ry, rz;
        read_positions(catch_parameter("SPS", par), sx, sy, sz);
        read_positions(catch_parameter("RPS", par), rx, ry, rz);
        const int ns = (int)sx.size(), nr = (int)rx.size();
        char us[64];
        std::snprintf(us, sizeof us, "%.0f", std::stod(dts) * 1e6);
        std::vector<float> data((size_t)nt * nr);
        for (int s = 0; s < ns; s++) {
            const std::string name = "seismogram_nt" + std::to_string(nt) + "_nr" + std::to_string(nr) + "_" + us +
                                     "us_shot_" + std::to_string(s + 1) + ".bin";
            import_binary_float(folder + name, data.data(), (long)data.size());
            rtm_input(data.data(), nt, nr, dt, sx[s], sy[s], rx.data(), ry.data());
            export_binary_float(folder + "input_RTM_" + name, data.data(), (long)data.size());
        }
        std::cout << ns << " shots written to " << folder << "input_RTM_*" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    return 0;
}
