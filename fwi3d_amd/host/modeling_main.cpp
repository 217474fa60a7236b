// modeling_main -- forward modeling driver (src/modeling_main.cpp:3-29).
// Usage: modeling_main parameters.txt
#include <chrono>
#include <iostream>

#include "modeling.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt\n";
        return 1;
    }
    try {
        Modeling modeling;
        modeling.parameters = std::string(argv[1]);
        modeling.set_parameters();
        auto ti = std::chrono::system_clock::now();
        for (modeling.srcId = 0; modeling.srcId < modeling.geometry->nsrc;
             modeling.srcId += modeling.shots_per_launch()) {
            modeling.show_information();
            modeling.initialization();
            modeling.forward_solver();
            modeling.get_seismogram();
        }
        auto tf = std::chrono::system_clock::now();
        std::chrono::duration<double> elapsed_seconds = tf - ti;
        std::cout << "\nRun time: " << elapsed_seconds.count() << " s." << std::endl;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    return 0;
}
