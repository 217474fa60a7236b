// modeling.cpp -- see modeling.hpp.
#include "modeling.hpp"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>

#include "admin.hpp"

namespace {
constexpr int WIDTH = 80;
}

Modeling::~Modeling() {
    if (ctx) fwi_destroy(ctx);
    delete geometry;
}

void Modeling::check(int rc, const std::string& what) const {
    if (rc != FWI_OK)
        throw std::runtime_error(what + " failed (" + std::to_string(rc) + "): " + fwi_last_error(ctx));
}

// modeling.cu:3-15
void Modeling::set_parameters() {
    title = "Seismic Modeling";
    set_main_parameters();
    set_build_parameters();
    set_abc_dampers();
    set_wavelet();
    set_geometry();
    set_properties();
    set_seismograms();
    create_context(nb, &ctx);
    check(fwi_set_dampers(ctx, b1d.data()), "fwi_set_dampers");
}

// modeling.cu:17-34
void Modeling::set_main_parameters() {
    nx = std::stoi(catch_parameter("x_samples", parameters));
    ny = std::stoi(catch_parameter("y_samples", parameters));
    nz = std::stoi(catch_parameter("z_samples", parameters));
    dh = std::stof(catch_parameter("model_spacing", parameters));
    nt = std::stoi(catch_parameter("time_samples", parameters));
    dt = std::stof(catch_parameter("time_spacing", parameters));
    fmax = std::stof(catch_parameter("max_frequency", parameters));
    data_folder = catch_parameter("mod_output_folder", parameters);
    idh2 = 1.0f / (dh * dh);
    idh3 = 1.0f / (dh * dh * dh);
}

void Modeling::set_build_parameters() {
    batch = std::max(1, std::stoi(catch_parameter_or("batch_shots", parameters, "1")));
    rbc_seed = std::stoull(catch_parameter_or("rbc_seed", parameters, "20261203"));
    rbc_amplitude = std::stof(catch_parameter_or("rbc_amplitude", parameters, "0.6"));
    const std::string style = catch_parameter_or("rbc_style", parameters, "blobs");
    if (style == "blobs") rbc_blobs = true;
    else if (style == "ramp") rbc_blobs = false;
    else throw std::invalid_argument("rbc_style must be 'blobs' or 'ramp', not '" + style + "'");
    const std::string a = catch_parameter_or("arithmetic", parameters, "reference");
    if (a == "reference") arith = FWI_ARITH_REFERENCE;
    else if (a == "fp32") arith = FWI_ARITH_FP32;
    else throw std::invalid_argument("arithmetic must be 'reference' or 'fp32'");
    const std::string at = catch_parameter_or("autotune", parameters, "on");
    if (at == "off") autotune = 0;
    else if (at == "on") autotune = 1;
    else if (at == "all") autotune = 2;
    else throw std::invalid_argument("autotune must be 'off', 'on' or 'all', not '" + at + "'");
    autotune_cache = catch_parameter_or("autotune_cache", parameters, "");
    physics = 0;
    std::string corr = catch_parameter_or("backward_corrections", parameters, "none");
    size_t p0 = 0;
    while (p0 <= corr.size()) {
        size_t p1 = corr.find(',', p0);
        if (p1 == std::string::npos) p1 = corr.size();
        std::string tok = corr.substr(p0, p1 - p0);
        tok.erase(0, tok.find_first_not_of(" \t"));
        tok.erase(tok.find_last_not_of(" \t") + 1);
        if (tok == "gradient_dyy") physics |= FWI_PHYS_GRAD_DYY;
        else if (tok == "subtract_source") physics |= FWI_PHYS_SUBTRACT_SOURCE;
        else if (tok == "aligned_imaging") physics |= FWI_PHYS_ALIGNED_IMAGING;
        else if (!(tok.empty() || tok == "none"))
            throw std::invalid_argument("backward_corrections: unknown item '" + tok + "'");
        p0 = p1 + 1;
    }
}

// modeling.cu:36-59: Ricker wavelet with peak frequency fmax/3.
void Modeling::set_wavelet() {
    wavelet.assign(nt, 0.0f);
    const float t0 = 2.0f * sqrtf((float)M_PI) / fmax;
    const float fc = fmax / (3.0f * sqrtf((float)M_PI));
    tlag = (int)(t0 / dt) + 1;
    for (int n = 0; n < nt; n++) {
        const float td = n * dt - t0;
        const float arg = (float)(M_PI * M_PI * M_PI * fc * fc * td * td);
        wavelet[n] = 1e5f * (1.0f - 2.0f * arg) * expf(-arg);
    }
}

// modeling.cu:61-70
void Modeling::set_geometry() {
    geometry = new Geometry();
    geometry->parameters = parameters;
    geometry->set_parameters();
}

// modeling.cu:72-81 (the device seismogram lives in the context)
void Modeling::set_seismograms() {
    data_folder = catch_parameter("mod_output_folder", parameters);
    seismogram.assign((size_t)nt * geometry->nrec, 0.0f);
}

// modeling.cu:83-140: only the 1-D table is needed; the library derives the
// 2-D / 3-D dampers with the reference's float operations.
void Modeling::set_abc_dampers() {
    ABC = true;
    abc_length = std::stof(catch_parameter("abc_length", parameters));
    abc_factor = std::stof(catch_parameter("abc_factor", parameters));
    nb = (int)(abc_length / dh) + 1;
    set_b1d();
}

// modeling.cu:96-99: the 1-D Cerjan table of nb cells.
void Modeling::set_b1d() {
    b1d.assign(nb, 0.0f);
    for (int i = 0; i < nb; i++) b1d[i] = expf(-powf(abc_factor * (nb - i), 2.0f));
}

// modeling.cu:142-179 (the edge padding itself runs on the device)
void Modeling::set_properties() {
    nPoints = (long)nx * ny * nz;
    model.assign(nPoints, 0.0f);
    import_binary_float(catch_parameter("model_file", parameters), model.data(), nPoints);
    vmax = *std::max_element(model.begin(), model.end());
    vmin = *std::min_element(model.begin(), model.end());
}

void Modeling::create_context(int boundary_cells, fwi_ctx** out) {
    nxx = nx + 2 * boundary_cells;
    nyy = ny + 2 * boundary_cells;
    nzz = nz + 2 * boundary_cells;
    volsize = (long)nxx * nyy * nzz;
    fwi_grid g{nxx, nyy, nzz, boundary_cells, dh, dt, nt, tlag};
    const int slots = std::min(batch, std::max(1, geometry->nsrc));
    batch = slots;
    int rc = fwi_create_batch(&g, slots, 0, out);
    if (rc != FWI_OK) throw std::runtime_error(std::string("fwi_create_batch: ") + fwi_last_error(nullptr));
    fwi_ctx* c = *out;
    check(fwi_set_arith(c, arith), "fwi_set_arith");
    check(fwi_set_physics(c, physics), "fwi_set_physics");
    check(fwi_set_model_interior(c, model.data()), "fwi_set_model_interior");
    check(fwi_set_wavelet(c, wavelet.data()), "fwi_set_wavelet");
}

float Modeling::cfl_velocity() const { return 0.45f * dh / dt; }

// The autotune key (not a reference parameter): launch shapes timed by
// fwi_autotune before a context's first pass of a kind.  Forward results do
// not depend on the shape; backward image / gradient sums do, through the
// grouping of shots, so "on" fixes the backward shape instead.
void Modeling::tune(fwi_ctx* c, int context, int kind) {
    const unsigned bit = 1u << (kind + 4 * context);
    if (autotune == 0 || (tuned_kinds & bit)) return;
    tuned_kinds |= bit;
    if (autotune == 1 && kind >= 2) {
        check(fwi_set_tuning(c, kind, 16, 1, 1), "fwi_set_tuning");
        return;
    }
    check(fwi_autotune(c, kind, 3, autotune_cache.empty() ? nullptr : autotune_cache.c_str()), "fwi_autotune");
}

// modeling.cu:558-592 (blobs) or the ramp boundary, seeded by (rbc_seed, shot)
void Modeling::random_boundary(fwi_ctx* c, int slot, int shot, float ratio, float varVp) {
    if (rbc_blobs)
        check(fwi_shot_random_boundary_blobs(c, slot, rbc_seed, shot, ratio, varVp, cfl_velocity()),
              "fwi_shot_random_boundary_blobs");
    else
        check(fwi_shot_random_boundary(c, slot, rbc_seed, shot, rbc_amplitude, cfl_velocity()),
              "fwi_shot_random_boundary");
}

// modeling.cu:270-272, 286-288
void Modeling::grid_index(float x, float y, float z, int* ix, int* iy, int* iz) const {
    *ix = (int)((x + 0.5f * dh) / dh) + nb;
    *iy = (int)((y + 0.5f * dh) / dh) + nb;
    *iz = (int)((z + 0.5f * dh) / dh) + nb;
}

// modeling.cu:264-300 for the shots [srcId, srcId + shot_count).
void Modeling::initialization() {
    first_shot = srcId;
    shot_count = std::min(batch, geometry->nsrc - srcId);
    std::vector<int> rx(geometry->nrec), ry(geometry->nrec), rz(geometry->nrec);
    for (int r = 0; r < geometry->nrec; r++)
        grid_index(geometry->xrec[r], geometry->yrec[r], geometry->zrec[r], &rx[r], &ry[r], &rz[r]);
    for (int slot = 0; slot < batch; slot++) {
        if (slot >= shot_count) {  // the last batch's unused slots carry no shot
            check(fwi_shot_disable(ctx, slot), "fwi_shot_disable");
            continue;
        }
        const int shot = first_shot + slot;
        grid_index(geometry->xsrc[shot], geometry->ysrc[shot], geometry->zsrc[shot], &sIdx, &sIdy, &sIdz);
        check(fwi_shot_set_source(ctx, slot, sIdx, sIdy, sIdz), "fwi_shot_set_source");
        check(fwi_shot_set_receivers(ctx, slot, geometry->nrec, rx.data(), ry.data(), rz.data()),
              "fwi_shot_set_receivers");
    }
}

// modeling.cu:309-322
void Modeling::forward_solver() {
    tune(ctx, 0, ABC ? 1 : 0);
    check(fwi_forward(ctx, ABC ? FWI_BOUNDARY_ABC : FWI_BOUNDARY_RBC), "fwi_forward");
}

// modeling.cu:302-307
void Modeling::get_seismogram() {
    for (int slot = 0; slot < shot_count; slot++) {
        check(fwi_shot_get_seismogram(ctx, slot, seismogram.data()), "fwi_shot_get_seismogram");
        const std::string data_file = data_folder + "seismogram_nt" + std::to_string(nt) + "_nr" +
                                      std::to_string(geometry->nrec) + "_" + std::to_string(int(1e6f * dt)) +
                                      "us_shot_" + std::to_string(first_shot + slot + 1) + ".bin";
        export_binary_float(data_file, seismogram.data(), (long)nt * geometry->nrec);
    }
}

// modeling.cu:243-262 (without clearing the terminal)
void Modeling::show_information() {
    const std::string line(WIDTH, '-');
    std::cout << line << '\n' << std::string((WIDTH - title.length()) / 2, ' ') << title << '\n' << line << '\n';
    std::cout << "Model dimensions: (z = " << (nz - 1) * dh << ", x = " << (nx - 1) * dh << ", y = " << (ny - 1) * dh
              << ") m\n\n";
    const int last = std::min(srcId + batch, geometry->nsrc);
    if (last - srcId > 1)
        std::cout << "Running shots " << srcId + 1 << "-" << last << " of " << geometry->nsrc << " in total\n\n";
    else
        std::cout << "Running shot " << srcId + 1 << " of " << geometry->nsrc << " in total\n\n";
    std::cout << "Current shot position: (z = " << geometry->zsrc[srcId] << ", x = " << geometry->xsrc[srcId]
              << ", y = " << geometry->ysrc[srcId] << ") m\n\n";
}
