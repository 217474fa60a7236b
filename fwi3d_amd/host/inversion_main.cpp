// inversion_main -- full-waveform inversion driver (src/inversion_main.cpp:3-32).
// Usage: inversion_main parameters.txt
#include <chrono>
#include <iostream>

#include "inversion.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt\n";
        return 1;
    }
    try {
        Inversion inversion;
        inversion.parameters = std::string(argv[1]);
        inversion.set_parameters();
        auto ti = std::chrono::system_clock::now();
        while (true) {
            inversion.compute_gradient();
            inversion.check_convergence();
            if (inversion.converged) break;
            inversion.optimization();
            inversion.update_model();
        }
        inversion.export_convergence();
        inversion.export_final_model();
        auto tf = std::chrono::system_clock::now();
        std::chrono::duration<double> elapsed_seconds = tf - ti;
        std::cout << "\nRun time: " << elapsed_seconds.count() << " s." << std::endl;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    return 0;
}
