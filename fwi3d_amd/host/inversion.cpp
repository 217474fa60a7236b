// inversion.cpp -- see inversion.hpp.
#include "inversion.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "admin.hpp"

Inversion::~Inversion() {
    if (rbc) fwi_destroy(rbc);
}

// inversion.cu:3-53
void Inversion::set_parameters() {
    title = "Full Waveform Inversion";
    set_main_parameters();
    set_build_parameters();
    set_abc_dampers();
    abc_nb = nb;
    set_wavelet();
    set_geometry();
    set_properties();
    set_seismograms();

    rbc_ratio = std::stof(catch_parameter("inv_rbc_ratio", parameters));
    rbc_varVp = std::stof(catch_parameter("inv_rbc_varVp", parameters));
    rbc_length = std::stof(catch_parameter("inv_rbc_length", parameters));
    rbc_nb = (int)(rbc_length / dh) + 1;

    zmask = std::stof(catch_parameter("depth_mask", parameters));
    input_folder = catch_parameter("inv_input_folder", parameters);
    input_prefix = catch_parameter("inv_input_prefix", parameters);
    output_folder = catch_parameter("inv_output_folder", parameters);
    residuo_folder = catch_parameter("inv_residuo_folder", parameters);
    max_iteration = std::stoi(catch_parameter("max_iteration", parameters));
    iteration = 0;
    const std::string upload = catch_parameter_or("fwi_model_upload", parameters, "updated");
    if (upload == "updated") upload_updated = true;
    else if (upload == "last_trial") upload_updated = false;
    else throw std::invalid_argument("fwi_model_upload must be 'updated' or 'last_trial', not '" + upload + "'");
    const std::string exg = catch_parameter_or("inv_export_gradient", parameters, "false");
    if (exg == "true") export_gradients = true;
    else if (exg == "false") export_gradients = false;
    else throw std::invalid_argument("inv_export_gradient must be 'true' or 'false', not '" + exg + "'");

    create_context(rbc_nb, &rbc);
    create_context(abc_nb, &ctx);  // leaves nb / nxx / nyy / nzz at the ABC grid
    check(fwi_set_dampers(ctx, b1d.data()), "fwi_set_dampers");

    obs_data.assign((size_t)nt * geometry->nrec, 0.0f);
    residual.assign((size_t)nt * geometry->nrec, 0.0f);
    gradient.assign(nPoints, 0.0f);
    trial.assign(nPoints, 0.0f);
    set_initial_model();
}

// inversion.cu:105-113: slowness squared of the starting model.
void Inversion::set_initial_model() {
    m.assign(nPoints, 0.0f);
    for (long q = 0; q < nPoints; q++) m[q] = 1.0f / (model[q] * model[q]);
}

// inversion.cu:172-176
void Inversion::set_obs_data(int shot) {
    const std::string data_file = input_folder + input_prefix + std::to_string(shot + 1) + ".bin";
    import_binary_float(data_file, obs_data.data(), (long)nt * geometry->nrec);
}

// Modeling::initialization on context c (boundary cells nbc) for the shots
// [first, first + count); the unused slots of a partial batch are disabled
// (no source, no receivers), so they add nothing to gradient or illumination.
void Inversion::place_shots(fwi_ctx* c, int nbc, int first, int count) {
    auto idx = [&](float v) { return (int)((v + 0.5f * dh) / dh) + nbc; };
    std::vector<int> rx(geometry->nrec), ry(geometry->nrec), rz(geometry->nrec);
    for (int r = 0; r < geometry->nrec; r++) {
        rx[r] = idx(geometry->xrec[r]);
        ry[r] = idx(geometry->yrec[r]);
        rz[r] = idx(geometry->zrec[r]);
    }
    for (int slot = 0; slot < batch; slot++) {
        if (slot >= count) {
            check(fwi_shot_disable(c, slot), "fwi_shot_disable");
            continue;
        }
        const int shot = first + slot;
        check(fwi_shot_set_source(c, slot, idx(geometry->xsrc[shot]), idx(geometry->ysrc[shot]),
                                  idx(geometry->zsrc[shot])),
              "fwi_shot_set_source");
        check(fwi_shot_set_receivers(c, slot, geometry->nrec, rx.data(), ry.data(), rz.data()),
              "fwi_shot_set_receivers");
    }
}

// inversion.cu:115-170 over batches of shots.
void Inversion::compute_gradient() {
    sum_res = 0.0;
    const bool final_check = iteration == max_iteration;  // last pass: misfit only
    if (!upload_updated && !last_trial.empty())
        trial = last_trial;  // the device model the reference still holds (inversion.cu:206, :324-326)
    else
        for (long q = 0; q < nPoints; q++) trial[q] = 1.0f / sqrtf(m[q]);
    check(fwi_set_model_interior(ctx, trial.data()), "fwi_set_model_interior");
    if (!final_check) {
        check(fwi_set_model_interior(rbc, trial.data()), "fwi_set_model_interior");
        check(fwi_zero_volume(rbc, FWI_VOL_GRAD), "fwi_zero_volume");
        check(fwi_zero_volume(rbc, FWI_VOL_ILLUM), "fwi_zero_volume");
    }
    for (srcId = 0; srcId < geometry->nsrc; srcId += batch) {
        const int count = std::min(batch, geometry->nsrc - srcId);
        stage_info = "Computing synthetic data";
        show_inv_info();
        place_shots(ctx, abc_nb, srcId, count);
        tune(ctx, 0, 1);
        check(fwi_forward(ctx, FWI_BOUNDARY_ABC), "fwi_forward");  // get_cal_data, inversion.cu:178-188
        for (int slot = 0; slot < count; slot++) {                  // set_adjoint_source, :190-201
            set_obs_data(srcId + slot);
            double s = 0.0;
            check(fwi_shot_residual(ctx, slot, obs_data.data(), &s), "fwi_shot_residual");
            sum_res += s;
        }
        if (final_check) continue;
        stage_info = "Wavefield reconstruction: forward and backward propagation";
        show_inv_info();
        place_shots(rbc, rbc_nb, srcId, count);
        for (int slot = 0; slot < count; slot++) {  // update_RBC, :204-216
            random_boundary(rbc, slot, srcId + slot, rbc_ratio, rbc_varVp);
            check(fwi_shot_get_backward_data(ctx, slot, residual.data()), "fwi_shot_get_backward_data");
            check(fwi_shot_set_backward_data(rbc, slot, residual.data()), "fwi_shot_set_backward_data");
        }
        tune(rbc, 1, 0);
        tune(rbc, 1, 3);
        check(fwi_forward(rbc, FWI_BOUNDARY_RBC), "fwi_forward");  // :218-233
        check(fwi_grad_backward(rbc), "fwi_grad_backward");        // :235-253
    }
    if (!final_check) {
        float gmax = 0.0f;
        check(fwi_export_gradient(rbc, zmask, gradient.data(), &gmax), "fwi_export_gradient");  // :142-169
        if (export_gradients) {
            const std::string file = output_folder + "gradient_iteration_" + std::to_string(iteration + 1) + "_" +
                                     std::to_string(int(fmax)) + "Hz_" + std::to_string(nz) + "x" +
                                     std::to_string(nx) + "x" + std::to_string(ny) + "_" +
                                     std::to_string((int)(dh)) + "m.bin";
            export_binary_float(file, gradient.data(), nPoints);
        }
    }
}

// inversion.cu:280-289; the misfit is reported as sqrt(sum r^2) like the reference
void Inversion::check_convergence() {
    ++iteration;
    residuo.push_back((float)std::sqrt(sum_res));
    converged = iteration > max_iteration;
    if (converged) std::cout << "Final residuo: " << residuo.back() << "\n\n";
}

double Inversion::model_misfit(const std::vector<float>& velocity) {
    check(fwi_set_model_interior(ctx, velocity.data()), "fwi_set_model_interior");
    double total = 0.0;
    for (srcId = 0; srcId < geometry->nsrc; srcId += batch) {
        const int count = std::min(batch, geometry->nsrc - srcId);
        show_inv_info();
        place_shots(ctx, abc_nb, srcId, count);
        tune(ctx, 0, 1);
        check(fwi_forward(ctx, FWI_BOUNDARY_ABC), "fwi_forward");
        for (int slot = 0; slot < count; slot++) {
            set_obs_data(srcId + slot);
            double s = 0.0;
            check(fwi_shot_residual(ctx, slot, obs_data.data(), &s), "fwi_shot_residual");
            total += s;
        }
    }
    return total;
}

// inversion.cu:315-347
double Inversion::linesearch(float alpha) {
    for (long q = 0; q < nPoints; q++) {
        trial[q] = m[q] - alpha * gradient[q];
        trial[q] = 1.0f / sqrtf(trial[q]);
    }
    sum_res = model_misfit(trial);
    if (!upload_updated) last_trial = trial;
    return sum_res;
}

// inversion.cu:291-312 (num, den and the vertex as the reference writes them)
float parabola_vertex(float a0, float a1, float a2, float f0, float f1, float f2) {
    const float num = (a1 * a1 - a2 * a2) * f0 + (a2 * a2 - a0 * a0) * f1 + (a0 * a0 - a1 * a1) * f2;
    const float den = (a1 - a2) * f0 + (a2 - a0) * f1 + (a0 - a1) * f2;
    return 0.5f * (num / den);
}

// A vertex the reference would take but whose model the library cannot run:
// a non-finite step (three collinear misfits, den = 0), or a step that
// drives some slowness squared to zero or below (NaN velocities in the
// reference) or a velocity past the stability limit.
float usable_step(float vertex, float a1, float a2, float f1, float f2, const float* m, const float* g, long n,
                  float vcfl, bool* fell_back) {
    auto usable = [&](float a) {
        if (!std::isfinite(a)) return false;
        for (long q = 0; q < n; q++) {
            const float mm = m[q] - a * g[q];
            if (!(mm > 0.0f) || !(1.0f / sqrtf(mm) <= vcfl)) return false;
        }
        return true;
    };
    if (fell_back) *fell_back = false;
    if (usable(vertex)) return vertex;
    if (fell_back) *fell_back = true;
    const float better = f1 < f2 ? a1 : a2;
    return usable(better) ? better : 0.0f;
}

// inversion.cu:291-312: parabola through (0, f0), (a1, f1), (a2, f2).
void Inversion::optimization() {
    const float dm = (1.0f / (vmin * vmin)) - (1.0f / (vmax * vmax));
    const float a0 = 0.00f * dm, a1 = 0.02f * dm, a2 = 0.05f * dm;
    const float f0 = residuo.back();
    stage_info = "Optimization via parabolic linesearch: first modeling";
    const float f1 = (float)std::sqrt(linesearch(a1));
    stage_info = "Optimization via parabolic linesearch: final modeling";
    const float f2 = (float)std::sqrt(linesearch(a2));
    step = parabola_vertex(a0, a1, a2, f0, f1, f2);
    // %.9g prints every float exactly (it reads back to the same value)
    char line[256];
    std::snprintf(line, sizeof line, "Line search: a1 = %.9g, a2 = %.9g, f(0) = %.9g, f(a1) = %.9g, f(a2) = %.9g, "
                  "step = %.9g\n", a1, a2, f0, f1, f2, step);
    std::cout << line;
    // the reference's vertex stands unless its model cannot run; then the
    // better of the two trial steps (or no step) is taken
    bool fell_back = false;
    step = usable_step(step, a1, a2, f1, f2, m.data(), gradient.data(), nPoints, cfl_velocity(), &fell_back);
    if (fell_back) {
        std::snprintf(line, sizeof line,
                      "Line search: the vertex leaves the valid model range; step = %.9g (the better trial step)\n",
                      step);
        std::cout << line;
    }
}

// inversion.cu:349-362
void Inversion::update_model() {
    for (long q = 0; q < nPoints; q++) {
        m[q] = m[q] - step * gradient[q];
        trial[q] = 1.0f / sqrtf(m[q]);
    }
    const std::string file = output_folder + "model_FWI_iteration_" + std::to_string(iteration) + "_" +
                             std::to_string(int(fmax)) + "Hz_" + std::to_string(nz) + "x" + std::to_string(nx) + "x" +
                             std::to_string(ny) + "_" + std::to_string((int)(dh)) + "m.bin";
    export_binary_float(file, trial.data(), nPoints);
}

// inversion.cu:364-376
void Inversion::export_convergence() {
    const std::string path = residuo_folder + "convergence_" + std::to_string(max_iteration) + "_iterations.txt";
    std::ofstream resFile(path, std::ios::out);
    for (int r = 0; r <= max_iteration && r < (int)residuo.size(); r++) resFile << residuo[r] << "\n";
    std::cout << "Text file " << path << " was successfully written." << std::endl;
}

// inversion.cu:378-383
void Inversion::export_final_model() {
    for (long q = 0; q < nPoints; q++) trial[q] = 1.0f / sqrtf(m[q]);
    const std::string file = output_folder + "final_model_FWI_" + std::to_string(int(fmax)) + "Hz_" +
                             std::to_string(nz) + "x" + std::to_string(nx) + "x" + std::to_string(ny) + "_" +
                             std::to_string((int)(dh)) + "m.bin";
    export_binary_float(file, trial.data(), nPoints);
}

// inversion.cu:255-278 (without clearing the terminal)
void Inversion::show_inv_info() {
    show_information();
    const std::string line(80, '-');
    std::cout << line << "\n" << stage_info << "\n" << line << "\n\n";
    if (iteration > max_iteration)
        std::cout << "-------- Checking final residuo --------\n\n";
    else if (iteration == 0)
        std::cout << "-------- Computing first residuo --------\n";
    else
        std::cout << "-------- Computing iteration " << iteration << " of " << max_iteration
                  << " --------\n\nPrevious residuo: " << residuo.back() << "\n\n";
}
