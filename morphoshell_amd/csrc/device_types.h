// Types shared by the host context and the device kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ms {

// Per-shell scalars and status, resident in device memory (one entry per
// shell).  Kernels read them as block-uniform values through the chunk
// tables; the last block of each node launch does the per-step bookkeeping
// (buffer parity, step clock, dial clock, equilibrium verdicts).
struct ShellDev {
    // mechanics (calc_deformation_forces.cpp:93-95, calc_energy_densities.cpp:62-77)
    double nu;
    double gamma;          // bend_stiffening_scale_factor
    // dynamics (calc_non_deformation_forces.cpp:55-57, advance_dynamics.cpp:58-93)
    double dt;
    double eta;            // damping_factor
    double damp_den;       // ref_density * h_min * h_min
    double lim;            // 5e4 * char_force_scale
    double fchar;
    double cchar;
    // dialling clock (main.cpp:286-288, 478-488)
    double dial_last_eq;
    double dial_spacing;
    double dial_K;         // timesteps_per_dial_phase as double
    long long phase_steps;
    int dialling;
    int protocol;          // 1 blend, 3 LCE
    int from_ansatz;
    int gd;                // gradient-descent dynamics
    int has_a0;
    int has_b0;
    // ranges
    int node_begin, num_nodes;
    int tri_begin, num_tris;
    int node_chunk_begin, node_chunk_count;
    int tri_chunk_begin, tri_chunk_count;
    // state
    int par;               // which of the two x/v buffers holds the current state
    int halted;            // 1: skipped by ms_step
    int failed;            // set by a node launch that found |F| >= lim
    int at_eq;
    long long halt_step;   // step-clock value at which the shell halted
    long long fail_step;
    // last equilibrium check
    double max_f;
    double max_v;
    int f_node;
    int v_node;
    int checked;
    // Device-run relaxation protocol (main.cpp:269-507): status machine
    // dialling -> waiting_for_equil -> equil_reached, checks every
    // equil_interval steps while waiting, damping eta1 while dialling and
    // eta2 while waiting.  auto_protocol = 0 leaves the status to the host.
    int auto_protocol;
    int status;            // 0 dialling, 1 waiting_for_equil, 2 equil_reached (final)
    int check_due;         // the next step runs an equilibrium check
    int pause_on_eq;       // halt at every equilibrium (host handles the event)
    int num_eq;            // equilibria reached so far
    long long since_check;
    long long equil_interval;
    double eta1, eta2;
    double tol_f, tol_v;
    long long last_eq_step;
    double f_override;     // >= 0: dial factor of the next step (main.cpp:483-485 ordering)
    double f_targets;      // factor abar/bbar were last dialled at (do_dialling runs only
                           // while dialling or at step 0, main.cpp:334-353)
    long long own_steps;   // steps this shell has taken (halted shells do not count)
    // Per-shell uniform triangle data (bitwise equal over the shell's
    // triangles): the force kernels take them from here instead of streaming
    // the per-triangle arrays.  uni_b is set only for blend-protocol shells.
    long long num_logged;  // protocol events written to the check log (ring index = count % capacity)
    int uni_hmu;           // thickness and shear modulus uniform: h_u, mu_u
    int uni_b;             // bbar_info uniform: b_u
    double h_u, mu_u;
    double b_u[3];
    // LCE shells: the fused step keeps each tile entry's dialled metric
    // lambda^2 t t + lambda^-2nu r r (before any ansatz blend) in the entry's
    // unused bbar-info record slots, and reuses it while the dial factor
    // stays lce_f and the tile records are generation lce_gen (-1: none).
    double lce_f;
    long long lce_gen;
};

// Dial factor a step uses (main.cpp:286-288, 483-485).
__host__ __device__ inline double shell_dial_factor(const ShellDev& s) {
    double f = s.f_override >= 0.0 ? s.f_override
                                   : s.dial_last_eq + s.dial_spacing * ((double)s.phase_steps / s.dial_K);
    return f >= 1.0 ? 1.0 : f;
}

// A contiguous range of triangles or nodes of one shell, processed by one
// block: {begin, end, shell, index in the context's full chunk table}.  A
// step launch may take a table of the running shells' chunks only; the index
// keeps per-chunk partial results at their place in the full table.
typedef int4 Chunk;

// One event of the device-run protocol, as the reference's log reports it
// (main.cpp:303, 396-407; is_at_equilibrium.cpp:89-90): kind 0 is an
// equilibrium check at the shell's own step `step` (its input state) with
// the dial factor used, the force and speed ratios and their first-argmax
// nodes, and the verdict; kind 1 is the end of a dialling phase (the shell
// starts waiting for equilibrium at dial factor `dial_factor` on step `step`).
struct CheckRec {
    long long step;
    double dial_factor;
    double max_force_ratio;
    double max_speed_ratio;
    int max_force_node;
    int max_speed_node;
    int at_equilibrium;
    int kind;
};

// Global clock of a context: steps taken.
struct ClockDev {
    long long steps;
    long long pad_;
};

}  // namespace ms
