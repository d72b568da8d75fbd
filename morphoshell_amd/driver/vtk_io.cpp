// Reference-format VTK reading and writing for the host driver; see vtk_io.h.

#include "vtk_io.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <chrono>
#include <iomanip>
#include <limits>
#include <stdexcept>

namespace msrun {
namespace {

constexpr std::streamsize kLine = std::numeric_limits<std::streamsize>::max();

void skip_line(std::istream& in) { in.ignore(kLine, '\n'); }
void skip_to_space(std::istream& in) { in.ignore(kLine, ' '); }

// The next whitespace-delimited word must be `want`.
void expect_word(std::istream& in, const char* want, const char* msg) {
    std::string w;
    in >> w;
    if (w != want) throw std::runtime_error(msg);
}

}  // namespace

RefMesh read_ref_vtk(const std::string& path, double uniform_thickness, double uniform_shear_modulus) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Error: Problem opening reference data file.");
    RefMesh m;
    for (int i = 0; i < 4; ++i) skip_line(in);  // header, title, ASCII, DATASET
    skip_to_space(in);                          // "POINTS "
    in >> m.num_nodes;
    if (!(m.num_nodes > 0))
        throw std::runtime_error("Error: Problem with (or before) line giving number of nodes in reference data file.");
    skip_line(in);
    const int N = m.num_nodes;
    m.pos.assign(3 * (size_t)N, 0.0);
    for (int n = 0; n < N; ++n) {
        in >> m.pos[n] >> m.pos[(size_t)N + n] >> m.pos[2 * (size_t)N + n];
        skip_line(in);
    }
    expect_word(in, "POLYGONS",
                "Error: Problem with reference data file, at or before 'POLYGONS' line. E.g. may have provided or "
                "stated wrong number of nodes.");
    in >> m.num_tris;
    if (!(m.num_tris > 0))
        throw std::runtime_error("Error: Problem with line giving number of triangles in reference data file.");
    skip_line(in);
    const int T = m.num_tris;
    m.tris.assign(3 * (size_t)T, 0);
    bool zero = false;
    for (int t = 0; t < T; ++t) {
        in.ignore(1);  // the vertex count "3"
        in >> m.tris[3 * (size_t)t] >> m.tris[3 * (size_t)t + 1] >> m.tris[3 * (size_t)t + 2];
        skip_line(in);
        for (int k = 0; k < 3; ++k) zero |= m.tris[3 * (size_t)t + k] == 0;
    }
    if (!zero)
        throw std::runtime_error(
            "Error: MorphoShell requires the triangulation you provide to use node ids starting from 0. It appears "
            "that the triangulation you provided did not do this, since it didn't contain a 0 anywhere.");
    expect_word(in, "CELL_DATA",
                "Error: Problem with reference data file, at or before 'CELL_DATA' line. E.g. you may have provided "
                "or stated wrong number of triangles.");
    skip_line(in);
    skip_line(in);  // FIELD line
    auto read_info = [&](const char* name, const char* msg, int32_t& cols, std::vector<double>& v) {
        expect_word(in, name, msg);
        in >> cols;
        skip_line(in);
        if (!(cols > 0)) throw std::runtime_error(msg);
        v.assign((size_t)T * cols, 0.0);
        for (int t = 0; t < T; ++t) {
            for (int p = 0; p < cols; ++p) in >> v[(size_t)t * cols + p];
            skip_line(in);
        }
    };
    read_info("abar_info", "Error: Problem with reference data file, at or before an abar_info line.", m.abar_cols,
              m.abar_info);
    read_info("bbar_info",
              "Error: Problem with reference data file, at or before a bbar_info line. E.g. you may have provided "
              "wrong number of abar_info lines.",
              m.bbar_cols, m.bbar_info);
    auto read_column = [&](const char* name, const char* msg, std::vector<double>& v, double uniform) {
        expect_word(in, name, msg);
        skip_line(in);
        v.assign(T, 0.0);
        for (int t = 0; t < T; ++t) {
            in >> v[t];
            skip_line(in);
            if (uniform > 0) v[t] = uniform;
        }
    };
    read_column("ref_thicknesses",
                "Error: Problem with reference data file, at or before a ref_thicknesses line. E.g. you may have "
                "provided wrong number of bbar_info lines.",
                m.thickness, uniform_thickness);
    read_column("ref_shear_moduli",
                "Error: Problem with reference data file, at or before a ref_shear_moduli line. E.g. you may have "
                "provided wrong number of ref_thicknesses lines.",
                m.shear_modulus, uniform_shear_modulus);
    expect_word(in, "tri_tags", "Error: Problem with reference data file, at or before tri_tags.");
    skip_line(in);
    m.tri_tags.assign(T, 0);
    for (int t = 0; t < T; ++t) {
        in >> m.tri_tags[t];
        skip_line(in);
    }
    expect_word(in, "POINT_DATA",
                "Error: Problem with reference data file, at or before 'POINT_DATA'. E.g. you may have provided wrong "
                "number of ref_thickness values.");
    skip_line(in);
    skip_line(in);  // FIELD line
    expect_word(in, "constraint_indicators",
                "Error: Problem with reference data file, at or before constraint_indicators.");
    skip_line(in);
    m.constrained.assign(N, 0);
    for (int n = 0; n < N; ++n) {
        int c = 0;
        in >> c;
        if (c != 0 && c != 1 && in) in.setstate(std::ios::failbit);  // a bool extraction accepts only 0 and 1
        m.constrained[n] = (uint8_t)c;
        skip_line(in);
    }
    expect_word(in, "node_tags", "Error: Problem with reference data file, at or before node_tags.");
    skip_line(in);
    m.node_tags.assign(N, 0);
    for (int n = 0; n < N; ++n) {
        in >> m.node_tags[n];
        skip_line(in);
    }
    if (in.bad() || in.fail())
        throw std::runtime_error(
            "Error: Unkown problem with reference data file. One of .bad(), .fail() was true - investigate! E.g. you "
            "may not have supplied enough node indicator data in the file, so the file may have reached its end "
            "prematurely.");
    std::string rest;
    in >> rest;
    if (!in.eof())
        throw std::runtime_error(
            "Error: Did not reach end of reference data file as expected. Check that it has exactly the correct "
            "format.");
    for (int t = 0; t < 3 * T; ++t)
        if (m.tris[t] < 0 || m.tris[t] >= N)
            throw std::runtime_error("Error: a triangle refers to a node id outside [0, number of nodes).");
    return m;
}

namespace {

constexpr char kMagic[8] = {'M', 'S', 'B', 'I', 'N', '0', '0', '1'};

struct SidecarKey {
    uint64_t size;
    int64_t mtime;
};

SidecarKey key_of(const std::string& path) {
    namespace fs = std::filesystem;
    const auto t = fs::last_write_time(path).time_since_epoch();
    return {(uint64_t)fs::file_size(path), (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t).count()};
}

template <class T>
void put(std::ofstream& o, const std::vector<T>& v) {
    const uint64_t n = v.size();
    o.write(reinterpret_cast<const char*>(&n), sizeof n);
    o.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(n * sizeof(T)));
}
template <class T>
bool get(std::ifstream& in, std::vector<T>& v, uint64_t expect) {
    uint64_t n = 0;
    in.read(reinterpret_cast<char*>(&n), sizeof n);
    if (!in || n != expect) return false;
    v.resize(n);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
    return (bool)in;
}

bool load_sidecar(const std::string& bin, const SidecarKey& k, RefMesh& m) {
    std::ifstream in(bin, std::ios::binary);
    if (!in) return false;
    char magic[8];
    SidecarKey kk{};
    in.read(magic, 8);
    in.read(reinterpret_cast<char*>(&kk), sizeof kk);
    if (!in || std::memcmp(magic, kMagic, 8) != 0 || kk.size != k.size || kk.mtime != k.mtime) return false;
    int32_t hdr[4];
    in.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    if (!in || hdr[0] <= 0 || hdr[1] <= 0 || hdr[2] <= 0 || hdr[3] <= 0) return false;
    m.num_nodes = hdr[0];
    m.num_tris = hdr[1];
    m.abar_cols = hdr[2];
    m.bbar_cols = hdr[3];
    const uint64_t N = (uint64_t)m.num_nodes, T = (uint64_t)m.num_tris;
    if (!(get(in, m.pos, 3 * N) && get(in, m.tris, 3 * T) && get(in, m.abar_info, T * m.abar_cols) &&
          get(in, m.bbar_info, T * m.bbar_cols) && get(in, m.thickness, T) && get(in, m.shear_modulus, T) &&
          get(in, m.tri_tags, T) && get(in, m.constrained, N) && get(in, m.node_tags, N)))
        return false;
    for (int32_t v : m.tris)
        if (v < 0 || v >= m.num_nodes) return false;
    for (uint8_t c : m.constrained)
        if (c > 1) return false;
    return true;
}

void save_sidecar(const std::string& bin, const SidecarKey& k, const RefMesh& m) {
    const std::string tmp = bin + ".tmp";
    {
        std::ofstream o(tmp, std::ios::binary);
        if (!o) return;
        o.write(kMagic, 8);
        o.write(reinterpret_cast<const char*>(&k), sizeof k);
        const int32_t hdr[4] = {m.num_nodes, m.num_tris, m.abar_cols, m.bbar_cols};
        o.write(reinterpret_cast<const char*>(hdr), sizeof hdr);
        put(o, m.pos);
        put(o, m.tris);
        put(o, m.abar_info);
        put(o, m.bbar_info);
        put(o, m.thickness);
        put(o, m.shear_modulus);
        put(o, m.tri_tags);
        put(o, m.constrained);
        put(o, m.node_tags);
        if (!o) return;
    }
    std::error_code ec;
    std::filesystem::rename(tmp, bin, ec);
}

}  // namespace

RefMesh read_ref_vtk_cached(const std::string& path, double uniform_thickness, double uniform_shear_modulus,
                            bool& from_sidecar) {
    from_sidecar = false;
    std::error_code ec;
    if (!std::filesystem::exists(path, ec)) throw std::runtime_error("Error: Problem opening reference data file.");
    const SidecarKey k = key_of(path);
    const std::string bin = path + ".msbin";
    RefMesh m;
    if (load_sidecar(bin, k, m)) {
        from_sidecar = true;
        for (auto& h : m.thickness)
            if (uniform_thickness > 0) h = uniform_thickness;
        for (auto& mu : m.shear_modulus)
            if (uniform_shear_modulus > 0) mu = uniform_shear_modulus;
        return m;
    }
    m = read_ref_vtk(path, -1.0, -1.0);  // the file's own values go into the sidecar
    save_sidecar(bin, k, m);
    for (auto& h : m.thickness)
        if (uniform_thickness > 0) h = uniform_thickness;
    for (auto& mu : m.shear_modulus)
        if (uniform_shear_modulus > 0) mu = uniform_shear_modulus;
    return m;
}

void read_ansatz_vtk(const std::string& path, int32_t num_nodes, std::vector<double>& pos, double& dial_factor) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Error: Problem opening ansatz data file.");
    skip_line(in);
    skip_to_space(in);  // "dial_factor "
    skip_to_space(in);  // "= "
    in >> dial_factor;
    skip_line(in);
    if (!(0 <= dial_factor) || !(dial_factor <= 1))
        throw std::runtime_error("Error: Problem reading ansatz data file. Remember it must have exactly the correct "
                                 "format.");
    skip_line(in);  // ASCII
    skip_line(in);  // DATASET
    skip_to_space(in);
    int32_t n_stated = 0;
    in >> n_stated;
    if (n_stated != num_nodes)
        throw std::runtime_error(
            "Error: Problem with (or before) line giving number of nodes in ansatz data file. E.g. you may have "
            "supplied ansatz and ref data files with inconsistent numbers of nodes.");
    skip_line(in);
    const size_t N = (size_t)num_nodes;
    pos.assign(3 * N, 0.0);
    for (size_t n = 0; n < N; ++n) {
        in >> pos[n] >> pos[N + n] >> pos[2 * N + n];
        skip_line(in);
    }
    if (in.bad() || in.fail())
        throw std::runtime_error(
            "Error: Unkown problem with ansatz data file. One of .bad(), .fail() was true - investigate!  E.g. you "
            "may not have supplied enough ansatz node positions, so the file reached its end prematurely.");
}

void write_output_vtk(const std::string& dir, int64_t stepcount, double dial_factor, const RefMesh& m,
                      const OutputFields& f) {
    const std::string path = dir + "/stepcount_" + std::to_string(stepcount) + "_output.vtk";
    std::ofstream out(path);
    if (!out) throw std::runtime_error("Error: Problem creating or opening vtk output file.");
    const int N = m.num_nodes, T = m.num_tris;
    out << std::scientific << std::setprecision(14);
    out << "# vtk DataFile Version 4.2\n"
        << "dial_factor = " << dial_factor << ", stepcount = " << stepcount << "\nASCII\nDATASET POLYDATA\n"
        << "POINTS " << N << " double\n";
    for (int n = 0; n < N; ++n)
        out << f.dofs[n] << " " << f.dofs[(size_t)N + n] << " " << f.dofs[2 * (size_t)N + n] << "\n";
    out << "POLYGONS " << T << " " << 4 * T << "\n";
    for (int t = 0; t < T; ++t)
        out << "3 " << m.tris[3 * (size_t)t] << " " << m.tris[3 * (size_t)t + 1] << " " << m.tris[3 * (size_t)t + 2]
            << "\n";
    out << "CELL_DATA " << T << "\n";
    auto scalars = [&](const char* name, const double* v, int count) {
        out << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
        for (int i = 0; i < count; ++i) out << v[i] << "\n";
    };
    scalars("gauss_curv", f.gauss_curv, T);
    scalars("mean_curv", f.mean_curv, T);
    scalars("dimless_stretch_energy_density", f.stretch_density, T);
    scalars("dimless_bend_energy_density", f.bend_density, T);
    scalars("strain_measure", f.strain, T);
    // write_output.cpp:136-156: the two kept eigenpairs of the Cauchy stress
    out << "FIELD cauchy_stress_info 4\n";
    auto field = [&](const char* name, int ncomp, int first) {
        out << name << " " << ncomp << " " << T << " double\n";
        for (int t = 0; t < T; ++t) {
            for (int c = 0; c < ncomp; ++c) out << (c ? " " : "") << f.stresses[(size_t)(first + c) * T + t];
            out << "\n";
        }
    };
    field("dimless_cauchy_stress_eigval_1", 1, 0);
    field("cauchy_stress_eigvec_1", 3, 1);
    field("dimless_cauchy_stress_eigval_2", 1, 4);
    field("cauchy_stress_eigvec_2", 3, 5);
    out << "POINT_DATA " << N << "\n";
    scalars("angleDeficit", f.angle_deficit, N);
    if (!out) throw std::runtime_error("Error: Problem writing vtk output file.");
}

}  // namespace msrun
