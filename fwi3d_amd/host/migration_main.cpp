// migration_main -- reverse time migration driver (src/migration_main.cpp:3-28).
// Usage: migration_main parameters.txt
#include <chrono>
#include <iostream>

#include "migration.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt\n";
        return 1;
    }
    try {
        Migration migration;
        migration.parameters = std::string(argv[1]);
        migration.set_parameters();
        auto ti = std::chrono::system_clock::now();
        for (migration.srcId = 0; migration.srcId < migration.geometry->nsrc;
             migration.srcId += migration.shots_per_launch()) {
            migration.forward_propagation();
            migration.backward_propagation();
        }
        auto tf = std::chrono::system_clock::now();
        migration.export_seismic();
        std::chrono::duration<double> elapsed_seconds = tf - ti;
        std::cout << "\nRun time: " << elapsed_seconds.count() << " s." << std::endl;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    return 0;
}
