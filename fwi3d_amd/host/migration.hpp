// migration.hpp -- reverse time migration (src/migration/migration.cuh:6-38):
// random-boundary forward pass, time-reversed source reconstruction with
// receiver back-propagation and cross-correlation imaging, then the
// illumination-compensated Laplacian image.
#pragma once

#include "modeling.hpp"

class Migration : public Modeling {
private:
    float rbc_ratio = 0.0f;
    float rbc_varVp = 0.0f;
    float rbc_length = 0.0f;

    std::vector<float> image;
    std::string stage_info;
    std::string input_folder;
    std::string input_prefix;
    std::string output_folder;

    void show_mig_info();
    void set_seismic_source();

public:
    void set_parameters();
    void forward_propagation();
    void backward_propagation();
    void export_seismic();
};
