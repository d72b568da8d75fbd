// Settings file of a run (the reference's format, read_settings.cpp:60-262):
// "key = value;" pairs, "//" comments to the end of the line, whitespace
// ignored, the reference's 23 keys first and in its order, then the keys this
// build appends.
#pragma once

#include <cstdint>
#include <string>

namespace msrun {

struct RunSettings {
    // the reference's keys, in file order
    double outputs_per_char_long_time = -1.0;
    double ref_thickness_if_uniform = -1.0;
    double timestep_prefactor = 0.6;
    bool dialling_from_ansatz_rather_than_ref_state = false;
    double dial_factors_to_hold_at_spacing = 1.0;
    double dial_phase_time_prefactor = 0.25;
    double time_between_equil_checks_prefactor = 0.01;
    double damping_prefactor_1 = 1.0;
    double damping_prefactor_2 = 1.0;
    int num_steps_gradient_descent = 0;  // num_steps_to_use_gradient_descent_dynamics_for_before_switching_to_newtonian_dynamics
    double force_to_char_force_ratio_equil_threshold = 1e-6;
    double speed_to_char_speed_ratio_equil_threshold = 1e-5;
    double def_poisson_ratio = 0.5;
    double ref_shear_modulus_if_uniform = -1.0;
    double ref_density = 1.0;
    double bend_stiffening_scale_factor = 1e17;
    double patch_mat_dimless_conditioning_thresh = 10.0;
    double slide_stiffness_prefactor = 0.0;
    double slide_friction_coefficent = 0.0;
    double slide_speed_prefactor = 0.0;
    bool using_initial_glass_slide_z_coords_from_settings_file = false;
    double initial_upper_slide_z_coord = 0.0;
    double initial_lower_slide_z_coord = 0.0;
    // appended keys
    std::string precision = "fp64";       // fp64 | mixed
    int dial_protocol = 1;                // 1 blend (do_dialling.cpp:83-84) | 3 LCE (104-124)
    int64_t timesteps_per_dial_phase = 0;  // > 0 replaces the derived dial-phase length
    int64_t max_steps = -1;               // > 0 stops the run after that many steps
};

// Throws std::runtime_error with the reference's messages.
RunSettings read_run_settings(const std::string& path, bool ansatz_given);

}  // namespace msrun
