// inversion.hpp -- full-waveform inversion (src/inversion/inversion.cuh:6-83).
//
// Two contexts, as the reference's dual padding (inversion.cu:55-103): an
// ABC-padded one for calculated data and line-search modeling, and an
// RBC-padded one for the random-boundary forward pass and the time-reversed
// gradient pass.  The misfit and the gradient normalisation are reduced in
// FP64 / exactly on the GPU.  Deliberate differences from the reference:
// the updated model is uploaded before the next iteration (inversion.cu:
// 349-362 never uploads it, so the reference computes the next gradient
// and misfit with the last line-search trial model; the key
// `fwi_model_upload = last_trial` restores that), and the illumination is
// cleared for every gradient (inversion.cu:51 never clears it).
#pragma once

#include "modeling.hpp"

// The parabolic line search's vertex (inversion.cu:291-312): the minimum of
// the parabola through (a0, f0), (a1, f1), (a2, f2), in float as the
// reference computes it (non-finite when the three misfits are collinear).
float parabola_vertex(float a0, float a1, float a2, float f0, float f1, float f2);

// The step the driver takes: the vertex when the model m - vertex * g keeps
// every slowness squared positive and every velocity 1/sqrt(m) within vcfl;
// otherwise the better trial step (a1 if f1 < f2, else a2) when its model
// does; otherwise 0.  *fell_back says whether the vertex was replaced.
float usable_step(float vertex, float a1, float a2, float f1, float f2, const float* m, const float* g, long n,
                  float vcfl, bool* fell_back);

class Inversion : public Modeling {
private:
    float step = 0.0f;
    float zmask = 0.0f;
    double sum_res = 0.0;

    int iteration = 0;
    int max_iteration = 0;

    float rbc_ratio = 0.0f;
    float rbc_varVp = 0.0f;
    float rbc_length = 0.0f;
    int abc_nb = 0, rbc_nb = 0;

    std::vector<float> m;         // slowness squared 1/v^2, interior
    std::vector<float> gradient;  // interior, normalised
    std::vector<float> trial;     // interior velocity of a line-search trial
    std::vector<float> last_trial;  // velocity of the latest trial (fwi_model_upload = last_trial)
    bool upload_updated = true;     // fwi_model_upload = updated (default) | last_trial
    bool export_gradients = false;  // inv_export_gradient = true: write each normalised gradient
    std::vector<float> obs_data;
    std::vector<float> residual;

    fwi_ctx* rbc = nullptr;

    std::string stage_info;
    std::string input_folder;
    std::string input_prefix;
    std::string output_folder;
    std::string residuo_folder;

    std::vector<float> residuo;

    void set_initial_model();
    void set_obs_data(int shot);
    void show_inv_info();
    void place_shots(fwi_ctx* c, int nbc, int first, int count);
    double model_misfit(const std::vector<float>& velocity);
    double linesearch(float alpha);

public:
    bool converged = false;

    ~Inversion() override;

    void set_parameters();

    void compute_gradient();
    void check_convergence();

    void optimization();
    void update_model();

    void export_final_model();
    void export_convergence();

    float last_step() const { return step; }
};
