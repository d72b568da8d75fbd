// Input and output files of a run, in the reference's formats.
//
// Input: the strict legacy-VTK reference mesh (read_vtk_data.cpp:84-371):
// POINTS, POLYGONS (ids from 0), CELL_DATA with abar_info, bbar_info,
// ref_thicknesses, ref_shear_moduli, tri_tags, POINT_DATA with
// constraint_indicators and node_tags, and nothing after them.  The optional
// ansatz file (read_vtk_data.cpp:380-445) is an earlier output: its second
// line carries the dial factor to start from, then POINTS.
// Output: stepcount_<n>_output.vtk per output (write_output.cpp:77-165).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace msrun {

struct RefMesh {
    int32_t num_nodes = 0, num_tris = 0;
    std::vector<double> pos;            // 3N, dofs SoA (x.., y.., z..)
    std::vector<int32_t> tris;          // 3T row-major
    int32_t abar_cols = 0, bbar_cols = 0;
    std::vector<double> abar_info;      // T x abar_cols row-major
    std::vector<double> bbar_info;      // T x bbar_cols row-major
    std::vector<double> thickness;      // T
    std::vector<double> shear_modulus;  // T
    std::vector<int32_t> tri_tags;      // T
    std::vector<uint8_t> constrained;   // N
    std::vector<int32_t> node_tags;     // N
};

// Throws std::runtime_error with the reference's messages.  Uniform
// thickness / shear modulus settings (> 0) replace the file's values.
RefMesh read_ref_vtk(const std::string& path, double uniform_thickness, double uniform_shear_modulus);

// The same through a binary sidecar <path>.msbin: loaded instead of parsing
// the text when its recorded size and modification time match the text
// file's, otherwise written after a parse (a failed write is not an error).
// from_sidecar tells which happened.
RefMesh read_ref_vtk_cached(const std::string& path, double uniform_thickness, double uniform_shear_modulus,
                            bool& from_sidecar);

// Ansatz positions (3N dofs SoA) and the dial factor to start from.
void read_ansatz_vtk(const std::string& path, int32_t num_nodes, std::vector<double>& pos, double& dial_factor);

struct OutputFields {
    const double* dofs;              // 3N
    const double* gauss_curv;        // T
    const double* mean_curv;         // T
    const double* stretch_density;   // T, already divided by the energy-density scale
    const double* bend_density;      // T, likewise
    const double* strain;            // T
    const double* stresses;          // 8T SoA: eigval_1 (over mu h), eigvec_1 x, y, z, eigval_2, eigvec_2 x, y, z
    const double* angle_deficit;     // N
};

void write_output_vtk(const std::string& dir, int64_t stepcount, double dial_factor, const RefMesh& m,
                      const OutputFields& f);


}  // namespace msrun
