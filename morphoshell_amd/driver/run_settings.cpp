// Settings reader of the host driver; see run_settings.h.

#include "run_settings.h"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <utility>
#include <vector>

namespace msrun {
namespace {

std::string strip(const std::string& raw) {
    std::string s = raw;
    std::replace(s.begin(), s.end(), '\r', '\n');
    s.push_back('\n');
    // a "//" comment runs to the end of its line, the newline included
    for (size_t p = s.find("//"); p != std::string::npos; p = s.find("//")) {
        const size_t e = s.find('\n', p);
        s.erase(p, e - p + 1);
    }
    s.erase(std::remove_if(s.begin(), s.end(), [](unsigned char c) { return std::isspace(c) != 0; }), s.end());
    return s;
}

bool to_bool(const std::string& v) {
    if (v == "true") return true;
    if (v == "false") return false;
    throw std::runtime_error("Error: you passed a std::string to boolstring_to_bool that read neither true nor false.");
}

}  // namespace

RunSettings read_run_settings(const std::string& path, bool ansatz_given) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Error: Problem opening settings file.");
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string s = strip(buf.str());
    std::vector<size_t> eq, sc;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '=') eq.push_back(i);
        if (s[i] == ';') sc.push_back(i);
    }
    if (eq.size() != sc.size())
        throw std::runtime_error(
            "Error: After removing all comments, the settings file did not have equal numbers of = and ; \n"
            " The most common mistake is to forget a semicolon.");
    for (size_t k = 0; k < eq.size(); ++k)
        if (eq[k] > sc[k] || (k + 1 < eq.size() && sc[k] > eq[k + 1]))
            throw std::runtime_error(
                "Error: After removing all comments, the = and ; in the settings file did not occur in the correct "
                "order.");
    std::vector<std::pair<std::string, std::string>> kv;
    size_t cur = 0;
    for (size_t k = 0; k < eq.size(); ++k) {
        kv.emplace_back(s.substr(cur, eq[k] - cur), s.substr(eq[k] + 1, sc[k] - eq[k] - 1));
        cur = sc[k] + 1;
    }

    RunSettings st;
    size_t j = 0;
    auto next = [&](const char* key) -> const std::string& {
        if (j >= kv.size() || kv[j].first != key)
            throw std::runtime_error(std::string("Error: something went wrong when reading the ") + key + " setting.");
        return kv[j++].second;
    };
    st.outputs_per_char_long_time = std::stod(next("outputs_per_char_long_time"));
    st.ref_thickness_if_uniform = std::stod(next("ref_thickness_if_uniform"));
    st.timestep_prefactor = std::stod(next("timestep_prefactor"));
    st.dialling_from_ansatz_rather_than_ref_state = to_bool(next("dialling_from_ansatz_rather_than_ref_state"));
    st.dial_factors_to_hold_at_spacing = std::stod(next("dial_factors_to_hold_at_spacing"));
    st.dial_phase_time_prefactor = std::stod(next("dial_phase_time_prefactor"));
    st.time_between_equil_checks_prefactor = std::stod(next("time_between_equil_checks_prefactor"));
    st.damping_prefactor_1 = std::stod(next("damping_prefactor_1"));
    st.damping_prefactor_2 = std::stod(next("damping_prefactor_2"));
    st.num_steps_gradient_descent =
        std::stoi(next("num_steps_to_use_gradient_descent_dynamics_for_before_switching_to_newtonian_dynamics"));
    st.force_to_char_force_ratio_equil_threshold = std::stod(next("force_to_char_force_ratio_equil_threshold"));
    st.speed_to_char_speed_ratio_equil_threshold = std::stod(next("speed_to_char_speed_ratio_equil_threshold"));
    st.def_poisson_ratio = std::stod(next("def_poisson_ratio"));
    st.ref_shear_modulus_if_uniform = std::stod(next("ref_shear_modulus_if_uniform"));
    st.ref_density = std::stod(next("ref_density"));
    st.bend_stiffening_scale_factor = std::stod(next("bend_stiffening_scale_factor"));
    st.patch_mat_dimless_conditioning_thresh = std::stod(next("patch_mat_dimless_conditioning_thresh"));
    st.slide_stiffness_prefactor = std::stod(next("slide_stiffness_prefactor"));
    st.slide_friction_coefficent = std::stod(next("slide_friction_coefficent"));
    st.slide_speed_prefactor = std::stod(next("slide_speed_prefactor"));
    st.using_initial_glass_slide_z_coords_from_settings_file =
        to_bool(next("using_initial_glass_slide_z_coords_from_settings_file"));
    st.initial_upper_slide_z_coord = std::stod(next("initial_upper_slide_z_coord"));
    st.initial_lower_slide_z_coord = std::stod(next("initial_lower_slide_z_coord"));
    for (; j < kv.size(); ++j) {
        const std::string& key = kv[j].first;
        const std::string& val = kv[j].second;
        if (key == "precision") {
            if (val != "fp64" && val != "mixed") throw std::runtime_error("Error: precision must be fp64 or mixed.");
            st.precision = val;
        } else if (key == "dial_protocol") {
            st.dial_protocol = std::stoi(val);
            if (st.dial_protocol != 1 && st.dial_protocol != 3)
                throw std::runtime_error("Error: dial_protocol must be 1 or 3.");
        } else if (key == "timesteps_per_dial_phase") {
            st.timesteps_per_dial_phase = std::stoll(val);
        } else if (key == "max_steps") {
            st.max_steps = std::stoll(val);
        } else {
            throw std::runtime_error("Error: unknown appended setting " + key + ".");
        }
    }
    if (st.dialling_from_ansatz_rather_than_ref_state && !ansatz_given)
        throw std::runtime_error(
            "Error: You cannot use the dialling_from_ansatz_rather_than_ref_state setting if you have not supplied an "
            "ansatz file.");
    if (st.dial_factors_to_hold_at_spacing > 1.0)
        throw std::runtime_error(
            "Error: The largest permitted value of dial_factors_to_hold_at_spacing is 1.0, because dial_factor never "
            "goes above 1.0, so anything larger would be meaningless.");
    if (st.slide_stiffness_prefactor > 0.0)
        throw std::runtime_error("Error: glass slides (slide_stiffness_prefactor > 0) are not supported by ms_run.");
    return st;
}

}  // namespace msrun
