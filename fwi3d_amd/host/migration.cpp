// migration.cpp -- see migration.hpp.
#include "migration.hpp"

#include <iostream>

#include "admin.hpp"

// migration.cu:3-36
void Migration::set_parameters() {
    title = "Reverse Time Migration";
    set_main_parameters();
    set_build_parameters();
    rbc_ratio = std::stof(catch_parameter("mig_rbc_ratio", parameters));
    rbc_varVp = std::stof(catch_parameter("mig_rbc_varVp", parameters));
    rbc_length = std::stof(catch_parameter("mig_rbc_length", parameters));
    nb = (int)(rbc_length / dh) + 1;
    ABC = false;
    set_wavelet();
    set_geometry();
    set_seismograms();
    set_properties();
    input_folder = catch_parameter("mig_input_folder", parameters);
    input_prefix = catch_parameter("mig_input_prefix", parameters);
    output_folder = catch_parameter("mig_output_folder", parameters);
    create_context(nb, &ctx);
    check(fwi_zero_volume(ctx, FWI_VOL_IMAGE), "fwi_zero_volume");
    check(fwi_zero_volume(ctx, FWI_VOL_ILLUM), "fwi_zero_volume");
    image.assign(nPoints, 0.0f);
}

void Migration::show_mig_info() {
    show_information();
    const std::string line(80, '-');
    std::cout << line << "\n" << stage_info << "\n" << line << "\n";
}

// migration.cu:49-59: a fresh random boundary per shot, then the forward pass.
void Migration::forward_propagation() {
    stage_info = "Forward propagation";
    show_mig_info();
    initialization();
    for (int slot = 0; slot < shot_count; slot++) random_boundary(ctx, slot, first_shot + slot, rbc_ratio, rbc_varVp);
    // the backward shape is chosen before the first forward pass: the
    // autotuner's trials overwrite the fields that pass leaves
    tune(ctx, 0, 2);
    forward_solver();
}

// migration.cu:61-81
void Migration::backward_propagation() {
    stage_info = "Backward propagation";
    show_mig_info();
    set_seismic_source();
    check(fwi_rtm_backward(ctx), "fwi_rtm_backward");
}

// migration.cu:83-88 (the unused slots of a last, partial batch are
// disabled: they have no receivers and inject nothing)
void Migration::set_seismic_source() {
    for (int slot = 0; slot < shot_count; slot++) {
        const std::string data_file = input_folder + input_prefix + std::to_string(first_shot + slot + 1) + ".bin";
        import_binary_float(data_file, seismogram.data(), (long)nt * geometry->nrec);
        check(fwi_shot_set_backward_data(ctx, slot, seismogram.data()), "fwi_shot_set_backward_data");
    }
}

// migration.cu:90-117; the illumination is written beside the image
void Migration::export_seismic() {
    check(fwi_export_image(ctx, image.data()), "fwi_export_image");
    const std::string dims = std::to_string(nz) + "x" + std::to_string(nx) + "x" + std::to_string(ny) + "_" +
                             std::to_string((int)(dh)) + "m.bin";
    export_binary_float(output_folder + "RTM_section_" + std::to_string(int(fmax)) + "Hz_" + dims, image.data(),
                        nPoints);
    std::vector<float> pad(volsize);
    check(fwi_get_volume(ctx, FWI_VOL_ILLUM, pad.data()), "fwi_get_volume");
    for (long q = 0; q < nPoints; q++) {
        const long k = q / ((long)nx * nz), j = (q - k * nx * nz) / nz, i = q - j * nz - k * nx * nz;
        image[q] = pad[(i + nb) + (j + nb) * nzz + (k + nb) * (long)nxx * nzz];
    }
    export_binary_float(output_folder + "RTM_illumination_" + std::to_string(int(fmax)) + "Hz_" + dims,
                        image.data(), nPoints);
}
