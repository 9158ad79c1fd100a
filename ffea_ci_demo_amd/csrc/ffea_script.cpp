// Host-side FFEA input reader for the blob path (see include/ffea_script.h).
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ffea_gpu.h"
#include "../../include/ffea_script.h"
#include "script_internal.h"

void ffea_internal_set_error(const std::string &msg);  // ffea_blob.hip

namespace {

// FFEA's internal units (include/dimensions.h:30-40)
namespace meso {
constexpr double length = 1.7e-10;
constexpr double Energy = 4.141945559999999e-21;
constexpr double mass = 1.994307387553024e-26;
constexpr double area = 2.89e-20;
constexpr double volume = 4.913e-30;
constexpr double force = 2.4364705882352941e-11;
constexpr double time = 3.7302915560886126e-13;
constexpr double pressure = 8.4305832688784826e8;
constexpr double pi = 3.141592653589793;
}  // namespace meso

struct Error {
    std::string msg;
};

[[noreturn]] void fail(const char *f, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    throw Error{buf};
}

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string dirname_of(const std::string &p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

std::string join(const std::string &dir, const std::string &f) {
    if (!f.empty() && f[0] == '/') return f;
    return dir + "/" + f;
}

// FFEA_input_reader::file_to_lines (FFEA_input_reader.cpp:38-114): drop <!-- --> comments
// (they may span lines), trim, keep non-empty lines
std::vector<std::string> file_to_lines(const std::string &path) {
    std::ifstream fin(path);
    if (!fin) fail("%s does not exist.", path.c_str());
    std::vector<std::string> out;
    std::string line;
    bool in_comment = false;
    while (std::getline(fin, line)) {
        std::string kept;
        size_t i = 0;
        while (i < line.size()) {
            if (in_comment) {
                const size_t e = line.find("-->", i);
                if (e == std::string::npos) {
                    i = line.size();
                } else {
                    in_comment = false;
                    i = e + 3;
                }
            } else {
                const size_t b = line.find("<!--", i);
                if (b == std::string::npos) {
                    kept += line.substr(i);
                    i = line.size();
                } else {
                    kept += line.substr(i, b - i);
                    in_comment = true;
                    i = b + 4;
                }
            }
        }
        kept = trim(kept);
        if (!kept.empty()) out.push_back(kept);
    }
    return out;
}

// the tag name of a block line "<name ...>"
std::string tag_name(const std::string &line) {
    std::string t = line;
    const size_t a = t.find('<');
    if (a != std::string::npos) t.erase(a, 1);
    const size_t b = t.rfind('>');
    if (b != std::string::npos) t.erase(b, 1);
    t = trim(t);
    const size_t sp = t.find(' ');
    return sp == std::string::npos ? t : t.substr(0, sp);
}

// FFEA_input_reader::extract_block (:116-169)
std::vector<std::string> extract_block(const std::string &title, int index, const std::vector<std::string> &in,
                                       bool mandatory = true) {
    int count = -1;
    bool copying = false;
    std::vector<std::string> out;
    for (const auto &l : in) {
        const std::string t = tag_name(l);
        if (t == title) {
            if (copying) fail("Shouldn't have found '%s' within %s block.", t.c_str(), title.c_str());
            ++count;
            if (count == index) {
                copying = true;
                continue;
            }
        }
        if (copying) {
            if (t == "/" + title) return out;
            out.push_back(l);
        }
    }
    if (copying) fail("Never found closing tag '%s'.", title.c_str());
    if (mandatory) fail("Specified block_index %d for block '%s' not found.", index, title.c_str());
    return {};
}

// FFEA_input_reader::parse_tag (:176-196): "<lvalue = rvalue>"
bool parse_tag(const std::string &line, std::string &lv, std::string &rv) {
    if (line.find('<') == std::string::npos) fail("Line '%s' is missing '<'\n", line.c_str());
    if (line.find('>') == std::string::npos) fail("Line '%s' is missing '>'\n", line.c_str());
    std::string t = line;
    t.erase(t.find('<'), 1);
    t.erase(t.rfind('>'), 1);
    t = trim(t);
    const size_t eq = t.find('=');
    if (eq == std::string::npos) {
        lv = t;
        rv.clear();
        return false;
    }
    lv = trim(t.substr(0, eq));
    rv = trim(t.substr(eq + 1));
    return true;
}

struct Params {
    ffea_script_params p{};
    int es_N[3] = {-1, -1, -1};
    double ssint_cutoff = 3e-9 / meso::length;
    // output and checkpoint files (SimulationParams.cpp:682-749, 972-990), as written: the
    // script's directory is prepended when the runner resolves them
    std::string traj_out, meas_out, det_out, ckpt_in, ckpt_out;
    bool hard_wall = false;  // some wall_* = HARD (SimulationParams.cpp:580-690)
};

// defaults of SimulationParams::SimulationParams (SimulationParams.cpp:26-117)
void set_defaults(Params &P) {
    ffea_script_params &p = P.p;
    p.dt = 1e-14 / meso::time;
    p.num_steps = (long long)1e11;
    p.check = 10000;
    p.kT = 4.11e-21 / meso::Energy;
    p.max_iterations_cg = 1000;
    p.epsilon2 = 0.01;
    p.rng_seed = (int)time(nullptr);
    p.calc_ssint = 1;
    p.calc_steric = 1;
    p.calc_noise = 1;
    p.calc_es = 0;
    p.calc_stokes = 1;
    p.stokes_visc = 1e-3 / (meso::pressure * meso::time);
    p.calc_kinetics = 0;
    p.calc_preComp = 0;
    p.calc_springs = 0;
    p.calc_ctforces = 0;
    p.move_into_box = 1;
    p.restart = -1;
    p.num_blobs = 0;
}

// SimulationParams::assign (SimulationParams.cpp:239-776): the keys of the blob path
// are read and converted; every other key the reference recognises is accepted and
// has no effect on this path; an unknown key is an error, as in the reference.
void assign(Params &P, const std::string &lv, const std::string &rv) {
    ffea_script_params &p = P.p;
    const char *r = rv.c_str();
    if (lv == "restart") p.restart = atoi(r);
    else if (lv == "dt") p.dt = atof(r) / meso::time;
    else if (lv == "epsilon") {
        const double e = atof(r);
        p.epsilon2 = e * e;
    } else if (lv == "num_steps") p.num_steps = (long long)atof(r);
    else if (lv == "max_iterations_cg") p.max_iterations_cg = atoi(r);
    else if (lv == "check") p.check = (int)atof(r);
    else if (lv == "num_blobs") p.num_blobs = atoi(r);
    else if (lv == "es_N_x") P.es_N[0] = atoi(r);
    else if (lv == "es_N_y") P.es_N[1] = atoi(r);
    else if (lv == "es_N_z") P.es_N[2] = atoi(r);
    else if (lv == "move_into_box") p.move_into_box = atoi(r);
    else if (lv == "rng_seed") p.rng_seed = (rv == "time") ? (int)time(nullptr) : atoi(r);
    else if (lv == "kT") p.kT = atof(r) / meso::Energy;
    else if (lv == "calc_ssint" || lv == "calc_vdw") p.calc_ssint = atoi(r);
    else if (lv == "calc_steric") p.calc_steric = atoi(r);
    else if (lv == "calc_es") p.calc_es = atoi(r);
    else if (lv == "calc_noise") p.calc_noise = atoi(r);
    else if (lv == "calc_preComp") p.calc_preComp = atoi(r);
    else if (lv == "calc_springs") p.calc_springs = atoi(r);
    else if (lv == "calc_kinetics") p.calc_kinetics = atoi(r);
    else if (lv == "calc_ctforces") p.calc_ctforces = atoi(r);
    else if (lv == "calc_stokes" || lv == "do_stokes") p.calc_stokes = atoi(r);
    else if (lv == "stokes_visc") p.stokes_visc = atof(r) / (meso::pressure * meso::time);
    else if (lv == "ssint_cutoff" || lv == "vdw_cutoff") P.ssint_cutoff = atof(r) / meso::length;
    else if (lv == "trajectory_out_fname") P.traj_out = rv;
    else if (lv == "measurement_out_fname") P.meas_out = rv;
    else if (lv == "det_measurement_out_fname") P.det_out = rv;
    else if (lv.compare(0, 5, "wall_") == 0 && lv.size() == 8 && (lv[5] == 'x' || lv[5] == 'y' || lv[5] == 'z') &&
             lv[6] == '_' && (lv[7] == '1' || lv[7] == '2'))
        P.hard_wall |= rv == "HARD";
    else if (lv == "checkpoint_in") P.ckpt_in = rv;
    else if (lv == "checkpoint_out") P.ckpt_out = rv;
    else {
        static const char *accepted[] = {
            "num_rods", "num_couplings", "num_conformations", "num_states", "es_update", "kinetics_update",
            "sticky_wall_xz", "es_h", "kappa", "dielec_ext", "epsilon_0", "calc_steric_rod", "calc_vdw_rod",
            "pbc_rod", "inc_self_ssint", "inc_self_vdw", "ssint_type", "vdw_type", "force_pbc", "steric_factor",
            "steric_dr", "flow_profile", "shear_rate", "flow_velocity", "trajectory_out_fname",
            "measurement_out_fname", "kinetics_out_fname", "vdw_in_fname", "ssint_in_fname",
            "vdw_forcefield_params", "ssint_forcefield_params", "ljparams", "lj_params", "rod_lj_in_fname",
            "rod_vdw_forcefield_params", "rod_lj_params", "checkpoint_in", "checkpoint_out", "beads_out_fname",
            "bsite_in_fname", "stress_out_fname", "vdw_steric_factor", "vdw_steric_dr", "trajbeads_fname",
            "beads_out_fname"};
        for (const char *a : accepted)
            if (lv == a) return;
        fail("'%s' is not a recognised lvalue\n", lv.c_str());
    }
}

struct BlobData {
    int nt = 0, ne = 0;
    int num_surface_nodes = 0;
    std::vector<double> pos;  // meso
    std::vector<int> top;
    std::vector<double> rho, A, B, G, E;
    std::vector<double> stokes_radius;  // meso
    std::vector<int> pinned;
    std::vector<int> ctf_node;
    std::vector<double> ctf_force;
    int num_surface_faces = 0;
    std::vector<int> surf_nodes;        // [faces][3] (.surf order, Face::n[0..2])
    std::vector<double> sf_force;       // linear surface constant forces: [sets][3] meso
    std::vector<int> sf_ptr{0}, sf_faces;
    std::vector<int> rf_node;           // rotational constant forces, list order ("all" expanded)
    std::vector<double> rf_force;       // [n] meso
    std::vector<double> rf_axis;        // [n][6] meso point, unit direction
    std::vector<char> rf_type;          // [n][2]
    int solver = FFEA_GPU_SOLVER_NOMASS_CG;  // <solver = ...> (World.cpp:2600-2625)
    double scale = 1.0;  // meso per file unit
    bool set_centroid = false;
    double centroid[3] = {0, 0, 0};
};

FILE *open_or_fail(const std::string &f) {
    FILE *in = fopen(f.c_str(), "r");
    if (!in) fail("Could not open file: %s", f.c_str());
    return in;
}

void expect_line(FILE *in, const char *a, const char *b, const char *what) {
    char line[256];
    if (!fgets(line, sizeof line, in)) fail("Error reading first line of %s file.", what);
    if (strcmp(line, a) != 0 && strcmp(line, b) != 0) fail("This is not a '%s' (read '%s').", b, line);
}

// Blob::load_nodes (Blob.cpp:1911-1992)
void load_nodes(BlobData &B, const std::string &f) {
    FILE *in = open_or_fail(f);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(in, fclose);
    expect_line(in, "walrus node file\n", "ffea node file\n", "node");
    int nn, ns, ni;
    if (fscanf(in, "num_nodes %d\n", &nn) != 1) fail("Error reading number of nodes.");
    if (fscanf(in, "num_surface_nodes %d\n", &ns) != 1) fail("Error reading number of surface nodes.");
    if (fscanf(in, "num_interior_nodes %d\n", &ni) != 1) fail("Error reading number of interior nodes.");
    char line[256];
    if (!fgets(line, sizeof line, in) || strcmp(line, "surface nodes:\n") != 0)
        fail("Could not find 'surface nodes:' line.");
    B.nt = nn;
    B.num_surface_nodes = ns;
    B.pos.resize((size_t)3 * nn);
    for (int i = 0; i < nn; ++i) {
        if (i == ns && (!fgets(line, sizeof line, in) || strcmp(line, "interior nodes:\n") != 0))
            fail("Could not find 'interior nodes:' line.");
        double x, y, z;
        if (fscanf(in, "%le %le %le\n", &x, &y, &z) != 3) fail("Error reading from nodes file at node %d", i);
        B.pos[3 * i] = B.scale * x;
        B.pos[3 * i + 1] = B.scale * y;
        B.pos[3 * i + 2] = B.scale * z;
    }
}

// Blob::load_topology (Blob.cpp:1996-2085)
void load_topology(BlobData &B, const std::string &f) {
    FILE *in = open_or_fail(f);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(in, fclose);
    expect_line(in, "walrus topology file\n", "ffea topology file\n", "topology");
    int ne, nse, nie;
    if (fscanf(in, "num_elements %d\n", &ne) != 1) fail("Error reading number of elements.");
    if (fscanf(in, "num_surface_elements %d\n", &nse) != 1) fail("Error reading number of surface elements.");
    if (fscanf(in, "num_interior_elements %d\n", &nie) != 1) fail("Error reading number of interior elements.");
    char line[256];
    if (!fgets(line, sizeof line, in) || strcmp(line, "surface elements:\n") != 0)
        fail("Could not find 'surface elements:' line.");
    B.ne = ne;
    B.top.resize((size_t)10 * ne);
    for (int i = 0; i < ne; ++i) {
        if (i == nse && (!fgets(line, sizeof line, in) || strcmp(line, "interior elements:\n") != 0))
            fail("Could not find 'interior elements:' line.");
        int *n = &B.top[(size_t)10 * i];
        if (fscanf(in, "%d %d %d %d %d %d %d %d %d %d\n", n, n + 1, n + 2, n + 3, n + 4, n + 5, n + 6, n + 7, n + 8,
                   n + 9) != 10)
            fail("Error reading from elements file at element %d", i);
        for (int k = 0; k < 10; ++k)
            if (n[k] < 0 || n[k] >= B.nt) fail("Error: Element %d references an out of bounds node index.", i);
    }
}

// Blob::load_surface (Blob.cpp:2090-2170): faces are validated and kept for surface
// constant forces
void load_surface(BlobData &B, const std::string &f) {
    FILE *in = open_or_fail(f);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(in, fclose);
    expect_line(in, "walrus surface file\n", "ffea surface file\n", "surface");
    int nf;
    if (fscanf(in, "num_surface_faces %d\n", &nf) != 1) fail("Error reading number of faces.");
    char line[256];
    if (!fgets(line, sizeof line, in) || strcmp(line, "faces:\n") != 0) fail("Could not find 'faces:' line.");
    // the midside sub-triangles of the four faces (slot sets; slots 4..9 on edges
    // 01 02 03 12 13 23): what tetra_element_linear::get_opposite_node accepts
    static const int faces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    static const int edge_slot[4][4] = {{-1, 4, 5, 6}, {4, -1, 7, 8}, {5, 7, -1, 9}, {6, 8, 9, -1}};
    std::vector<std::array<int, 3>> ok;
    for (auto &fc : faces) {
        const int a = fc[0], b = fc[1], c = fc[2];
        const int ab = edge_slot[a][b], bc = edge_slot[b][c], ca = edge_slot[c][a];
        for (auto t : {std::array<int, 3>{a, ab, ca}, {ab, b, bc}, {ca, bc, c}, {ab, bc, ca}}) {
            std::sort(t.begin(), t.end());
            ok.push_back(t);
        }
    }
    for (int i = 0; i < nf; ++i) {
        int el, n[3];
        if (fscanf(in, "%d %d %d %d\n", &el, n, n + 1, n + 2) != 4)
            fail("Error reading from surface file at face %d. There should be 4 space separated integers.", i);
        for (int k = 0; k < 3; ++k)
            if (n[k] < 0 || n[k] >= B.nt) fail("Error: Surface face %d references an out of bounds node index.", i);
        if (el < 0 || el >= B.ne) fail("Error: Surface face %d references an out of bounds element index.", i);
        std::array<int, 3> slots;
        for (int k = 0; k < 3; ++k) {
            int s = -1;
            for (int j = 0; j < 10; ++j)
                if (B.top[(size_t)10 * el + j] == n[k]) s = j;
            if (s < 0) fail("Specified node index does not belong to this element.");
            slots[k] = s;
        }
        std::sort(slots.begin(), slots.end());
        if (std::find(ok.begin(), ok.end(), slots) == ok.end()) fail("Error: Could not find the opposite node.");
        B.surf_nodes.insert(B.surf_nodes.end(), {n[0], n[1], n[2]});
    }
    B.num_surface_faces = nf;
}

// Blob::load_material_params (Blob.cpp:2241-2293)
void load_material(BlobData &B, const std::string &f) {
    FILE *in = open_or_fail(f);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(in, fclose);
    expect_line(in, "walrus material params file\n", "ffea material params file\n", "material params");
    int ne;
    if (fscanf(in, "num_elements %d\n", &ne) != 1) fail("Error reading number of elements.");
    if (ne != B.ne)
        fail("Number of elements in material params file (%d) does not match number of elements in topology file (%d).",
             ne, B.ne);
    B.rho.resize(ne);
    B.A.resize(ne);
    B.B.resize(ne);
    B.G.resize(ne);
    B.E.resize(ne);
    for (int i = 0; i < ne; ++i) {
        double density, sv, bv, sm, bm, diel;
        if (fscanf(in, "%le %le %le %le %le %le\n", &density, &sv, &bv, &sm, &bm, &diel) != 6)
            fail("Error reading from material params file at element %d. There should be 6 space separated real "
                 "values (density, shear_visc, bulk_visc, shear_mod, bulk_mod, dielectric).",
                 i);
        B.rho[i] = density * meso::volume / meso::mass;
        B.A[i] = sv / (meso::pressure * meso::time);
        B.B[i] = bv / (meso::pressure * meso::time) - (2.0 / 3.0) * sv;  // second coefficient, as read there
        B.G[i] = sm / meso::pressure;
        B.E[i] = bm / meso::pressure;
    }
}

// Blob::load_stokes_params (Blob.cpp:2297-2358)
void load_stokes(BlobData &B, const std::string &f) {
    FILE *in = open_or_fail(f);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(in, fclose);
    expect_line(in, "walrus stokes radii file\n", "ffea stokes radii file\n", "stokes radii");
    int nn;
    if (fscanf(in, "num_nodes %d\n", &nn) != 1) fail("Error reading number of nodes.");
    if (nn != B.nt)
        fail("Number of nodes in stokes radii file (%d) does not match number of nodes in nodes file (%d).", nn,
             B.nt);
    B.stokes_radius.resize(nn);
    for (int i = 0; i < nn; ++i) {
        double r;
        if (fscanf(in, "%le\n", &r) != 1)
            fail("Error reading from stokes radii file at node %d. There should be 1 real value (stokes radius).", i);
        B.stokes_radius[i] = r * B.scale;
        if (B.stokes_radius[i] < 1e-12 && B.stokes_radius[i] > 0.0)
            fail("Stokes Radius on node %d in this Blob is very small, %e.", i, B.stokes_radius[i]);
    }
}

// Blob::load_pinned_nodes (Blob.cpp:3024-3076)
void load_pins(BlobData &B, const std::string &f) {
    FILE *in = open_or_fail(f);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(in, fclose);
    expect_line(in, "walrus pinned nodes file\n", "ffea pinned nodes file\n", "pinned nodes");
    int np;
    if (fscanf(in, "num_pinned_nodes %d\n", &np) != 1) fail("Error reading number of pinned nodes.");
    char line[256];
    if (!fgets(line, sizeof line, in) || strcmp(line, "pinned nodes:\n") != 0)
        fail("Could not find 'pinned nodes:' line.");
    B.pinned.resize(np);
    for (int i = 0; i < np; ++i) {
        int k;
        if (fscanf(in, "%d\n", &k) != 1)
            fail("Error reading from pinned node file at face %d. There should be 1 integer per line.", i);
        if (k < 0 || k >= B.nt) fail("Error: Pinned node %d references an out of bounds node index.", i);
        B.pinned[i] = k;
    }
}

std::vector<std::string> split_ws(const std::string &s) {
    std::istringstream is(s);
    std::vector<std::string> out;
    std::string w;
    while (is >> w) out.push_back(w);
    return out;
}

// Blob::load_ctforces, linear node forces and linear surface forces (Blob.cpp:2478-2623,
// 2730-2806); rotational forces are refused
void load_ctforces(BlobData &B, int blob_index, const std::string &f) {
    std::vector<std::string> L = file_to_lines(f);
    if (L.size() < 4) fail("ctforces_fname '%s' is too short, something is wrong\n", f.c_str());
    int n_all = 0, n_lin = 0, n_rot = 0, n_surf = 0;
    size_t now = 1;
    for (; now < 5 && now < L.size(); ++now) {
        if (L[now].find(':') != std::string::npos) break;
        auto w = split_ws(L[now]);
        if (w.size() < 2) continue;
        if (w[0] == "num_ctforces") n_all = atoi(w[1].c_str());
        if (w[0] == "num_linear_forces") n_lin = atoi(w[1].c_str());
        if (w[0] == "num_rot_forces") n_rot = atoi(w[1].c_str());
        if (w[0] == "num_linear_surface_forces") n_surf = atoi(w[1].c_str());
    }
    size_t calc_lines = n_all + now + (n_lin ? 1 : 0) + (n_rot ? 1 : 0) + (n_surf ? 1 : 0);
    if (n_all != n_lin + n_rot + n_surf || n_all == 0 || calc_lines != L.size())
        fail("--- ABORTING. Something went wrong when parsing the header of the ctforces file\n"
             "--- Check it or turn calc_ctforces to 0 in the FFEA input file");
    if (n_lin > 0) {
        if (L[now].compare(0, 14, "linear forces:") == 0) ++now;
        else
            fail("Error reading the ctforces file; it should announce the start of linear ctforces, but instead "
                 "read: %s\n", L[now].c_str());
    }
    for (int i = 0; i < n_lin; ++i) {
        auto w = split_ws(L[now + i]);
        if (w.size() != 7) fail("Invalid line in the ctforces file:\n \t %s\n", L[now + i].c_str());
        if (atoi(w[4].c_str()) != blob_index || atoi(w[5].c_str()) != 0) continue;
        const double F = atof(w[0].c_str());
        double d[3] = {atof(w[1].c_str()), atof(w[2].c_str()), atof(w[3].c_str())};
        const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (nrm == 0) fail("Unable to normalize a vector with length zero");
        for (double &x : d) x /= nrm;  // normalize (mat_vec_fns_II.h:188-198)
        const double fx = F * d[0] / meso::force, fy = F * d[1] / meso::force, fz = F * d[2] / meso::force;
        if (w[6] == "all") {
            for (int j = 0; j < B.nt; ++j) {
                B.ctf_node.push_back(j);
                B.ctf_force.insert(B.ctf_force.end(), {fx, fy, fz});
            }
        } else {
            const int k = atoi(w[6].c_str());
            if (k < 0 || k >= B.nt) fail("constant force on an out of bounds node index %d", k);
            B.ctf_node.push_back(k);
            B.ctf_force.insert(B.ctf_force.end(), {fx, fy, fz});
        }
    }
    now += n_lin;
    if (n_rot > 0) {  // Blob.cpp:2630-2726
        if (now >= L.size() || L[now].compare(0, 18, "rotational forces:") != 0)
            fail("Wrong header; it should announce the start of rotational ctforces, but instead read: %s\n",
                 now < L.size() ? L[now].c_str() : "");
        ++now;
        const double mdfm1 = 1. / meso::force, mdlm1 = 1. / meso::length;
        for (int i = 0; i < n_rot; ++i) {
            auto w = split_ws(L[now + i]);
            if (w.size() != 11) fail("Invalid line in the ctforces file:\n \t %s\n", L[now + i].c_str());
            if (atoi(w[8].c_str()) != blob_index || atoi(w[9].c_str()) != 0) continue;
            const double F = atof(w[0].c_str()) * mdfm1;
            const std::string &type = w[1];
            double pt[3] = {atof(w[2].c_str()), atof(w[3].c_str()), atof(w[4].c_str())};
            double d[3] = {atof(w[5].c_str()), atof(w[6].c_str()), atof(w[7].c_str())};
            if (type.size() < 2 || (type[1] != 'f' && type[1] != 't'))
                fail("Invalid rotational force type: %s, in line read: %s\n", type.c_str(), L[now + i].c_str());
            if (type[0] == 'n')
                fail("rotational force with an axis given by two nodes ('n'): the reference's force loop never "
                     "forms r for it (Blob.cpp:1697-1706); give the axis by point and direction ('p')");
            if (type[0] != 'p')
                fail("Invalid rotational force: %s, in line read: %s\n", type.c_str(), L[now + i].c_str());
            const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (nrm == 0) fail("Unable to normalize a vector with length zero");
            for (double &x : d) x /= nrm;
            for (double &x : pt) x = mdlm1 * x;  // resize(mdlm1, ctf_p)
            auto add = [&](int node) {
                B.rf_node.push_back(node);
                B.rf_force.push_back(F);
                B.rf_axis.insert(B.rf_axis.end(), {pt[0], pt[1], pt[2], d[0], d[1], d[2]});
                B.rf_type.insert(B.rf_type.end(), {type[0], type[1]});
            };
            if (w[10] == "all") {
                for (int j = 0; j < B.nt; ++j) add(j);
            } else {
                const int k = atoi(w[10].c_str());
                if (k < 0 || k >= B.nt) fail("rotational force on an out of bounds node index %d", k);
                add(k);
            }
        }
        now += n_rot;
    }
    if (n_surf > 0) {  // Blob.cpp:2730-2806
        if (now >= L.size() || L[now].compare(0, 22, "linear surface forces:") != 0)
            fail("Wrong header; it should announce the start of the linear surface ctforces, but instead read: %s\n",
                 now < L.size() ? L[now].c_str() : "");
        ++now;
        for (int i = 0; i < n_surf; ++i) {
            auto w = split_ws(L[now + i]);
            if (w.size() < 7) fail("Invalid line in the ctforces file:\n \t %s\n", L[now + i].c_str());
            if (atoi(w[4].c_str()) != blob_index || atoi(w[5].c_str()) != 0) continue;
            if (w[6] == "all") fail("linear surface forces on 'all' faces: the reference's loader throws here");
            const double F = atof(w[0].c_str());
            double d[3] = {atof(w[1].c_str()), atof(w[2].c_str()), atof(w[3].c_str())};
            const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (nrm == 0) fail("Unable to normalize a vector with length zero");
            for (double &x : d) x /= nrm;
            B.sf_force.insert(B.sf_force.end(), {F * d[0] / meso::force, F * d[1] / meso::force, F * d[2] / meso::force});
            for (size_t j = 6; j < w.size(); ++j) {
                const int face = atoi(w[j].c_str());
                if (face < 0 || face >= B.num_surface_faces)
                    fail("linear surface force on an out of bounds face index %d", face);
                B.sf_faces.push_back(face);
            }
            B.sf_ptr.push_back((int)B.sf_faces.size());
        }
    }
}

}  // namespace

// one spring of the springs file (Spring.h; World::load_springs, World.cpp:3543-3643)
struct SpringData {
    double k = 0, l = 0;  // meso units
    int blob[2] = {0, 0}, conf[2] = {0, 0}, node[2] = {0, 0};
    bool active = false;  // World::activate_springs: both ends in the active (first) conformation
};

struct ffea_script {
    std::string path;
    Params P;
    std::vector<BlobData> blobs;
    std::vector<SpringData> springs;
    uint32_t rng_state[6] = {0, 0, 0, 0, 0, 0};
};

namespace {

void parse_vec3(const std::string &s, double *v) {
    std::string t;
    for (char c : s)
        if (c != '(' && c != ')') t += (c == ',') ? ' ' : c;
    std::istringstream is(t);
    is >> v[0] >> v[1] >> v[2];
    if (!is) fail("could not read a 3-vector from '%s'", s.c_str());
}

// World::load_springs (World.cpp:3543-3643), read with the reference's fscanf formats
void load_springs(ffea_script &S, const std::string &fname) {
    FILE *in = fopen(fname.c_str(), "rb");
    if (!in) fail("Failed to open springs file: '%s'", fname.c_str());
    char line[50] = {0};
    if (!fgets(line, sizeof line, in)) {
        fclose(in);
        fail("Error reading first line of spring file");
    }
    if (strcmp(line, "ffea springs file\n") != 0) {
        fclose(in);
        fail("This is not a 'ffea spring file' (read '%s')", line);
    }
    int num = 0;
    if (fscanf(in, "num_springs %d\n", &num) != 1) {
        fclose(in);
        fail("Error reading number of springs");
    }
    const int header = fscanf(in, "springs:\n");  // unchecked in the reference: the line may be absent
    (void)header;
    const int nb = (int)S.blobs.size();
    for (int i = 0; i < num; ++i) {
        SpringData sp;
        if (fscanf(in, "%lf %lf %d %d %d %d %d %d\n", &sp.k, &sp.l, &sp.blob[0], &sp.blob[1], &sp.conf[0], &sp.conf[1],
                   &sp.node[0], &sp.node[1]) != 8) {
            fclose(in);
            fail("Problem reading spring data from %s.Format is : \n\nffea spring file\nnum_springs ? \n"
                 "k l blob_index_0 blob_index 1 conformation_index_0 conformation_index_1 node_index_0 node_index_1\n",
                 fname.c_str());
        }
        sp.k *= meso::area / meso::Energy;
        sp.l /= meso::length;
        for (int j = 0; j < 2; ++j) {
            if (sp.blob[j] >= nb || sp.blob[j] < 0) {
                fclose(in);
                fail("In spring %d, blob index %d is out of bounds given the number of blobs defined (%d). Please fix. "
                     "Remember, indexing starts at ZERO!\n", i, j, nb);
            }
            if (sp.conf[j] >= 1 || sp.conf[j] < 0)
                printf("In spring %d, conformation index %d is out of bounds given the number of conformations defined "
                       "in blob %d (%d). Please fix. Remember, indexing starts at ZERO!\n", i, j, sp.blob[j], 1);
            if (sp.node[j] >= S.blobs[sp.blob[j]].nt || sp.node[j] < 0) {
                fclose(in);
                fail("In spring %d, node index %d is out of bounds given the number of nodes defined in blob %d, "
                     "conformation %d (%d). Please fix. Remember, indexing starts at ZERO!\n", i, j, sp.blob[j],
                     sp.conf[j], S.blobs[sp.blob[j]].nt);
            }
            if (sp.k < 0) {
                fclose(in);
                fail("In spring %d, spring constant, %e, < 0. This is not going to end well for you...\n", i, sp.k);
            }
            if (sp.l < 0) {
                fclose(in);
                fail("In spring %d, spring equilibrium length, %e, < 0. Reverse node definitions for consistency.\n",
                     i, sp.l);
            }
        }
        if (sp.blob[0] == sp.blob[1] && sp.conf[0] == sp.conf[1] && sp.node[0] == sp.node[1]) {
            fclose(in);
            fail("In spring %d, spring is connected to same node on same blob on same conformation. Will probably cause "
                 "an force calculation error.\n", i);
        }
        if (sp.blob[0] == sp.blob[1] && sp.conf[0] != sp.conf[1]) {
            fclose(in);
            fail("In spring %d, spring is connected two conformations of the same blob (blob_index = %d). This cannot "
                 "happen as conformations are mutually exclusive.\n", i, sp.blob[0]);
        }
        sp.active = sp.conf[0] == 0 && sp.conf[1] == 0;
        S.springs.push_back(sp);
    }
    fclose(in);
}

void load_script(ffea_script &S) {
    const std::string dir = dirname_of(S.path);
    std::vector<std::string> lines = file_to_lines(S.path);
    Params &P = S.P;
    set_defaults(P);
    std::vector<std::string> param = extract_block("param", 0, lines);
    for (const auto &l : param) {
        std::string lv, rv;
        parse_tag(l, lv, rv);
        assign(P, lv, rv);
    }
    ffea_script_params &p = P.p;
    // SimulationParams::validate, the checks on this path (SimulationParams.cpp:805-1060)
    if (p.num_blobs <= 0) fail("\tRequired: Number of Blobs, 'num_blobs', must be greater than 0.");
    if (p.dt <= 0) fail("Required: Timestep, 'dt', must be greater than 0.");
    if (p.max_iterations_cg <= 0) fail("Required: 'max_iterations_cg', must be greater than 0.");
    if (p.calc_noise != 0 && p.calc_noise != 1) fail("Required: 'calc_noise', must be 0 (no) or 1 (yes).");
    if (p.move_into_box != 0 && p.move_into_box != 1) fail("'move_into_box' must all be either 0 or 1.");
    if (!(p.calc_ssint || p.calc_steric || p.calc_es || p.calc_preComp)) P.es_N[0] = P.es_N[1] = P.es_N[2] = 0;

    std::vector<std::string> system = extract_block("system", 0, lines);
    std::string ctf_file;
    if (p.calc_ctforces == 1) {
        std::vector<std::string> inter = extract_block("interactions", 0, system);
        std::vector<std::string> ctf = extract_block("ctforces", 0, inter);
        if (ctf.size() != 1) fail("only a single field is allowed in <ctforces> block.");
        std::string lv, rv;
        parse_tag(ctf[0], lv, rv);
        if (lv != "ctforces_fname")
            fail("ctforces_fname not provided in <ctforces> block, while calc_ctforces set to 1.");
        ctf_file = join(dir, rv);
    }
    std::string springs_file;
    bool springs_given = false;
    if (p.calc_springs == 1) {  // World.cpp:2291-2305
        std::vector<std::string> inter = extract_block("interactions", 0, system);
        std::vector<std::string> spr = extract_block("springs", 0, inter);
        if (spr.size() > 1) fail("'Spring' block should only have 1 file.");
        if (spr.size() == 1) {
            std::string lv, rv;
            parse_tag(spr[0], lv, rv);
            springs_file = join(dir, rv);
            springs_given = true;
        }
    }
    for (int b = 0; b < p.num_blobs; ++b) {
        std::vector<std::string> blob = extract_block("blob", b, system);
        std::vector<std::string> conf = extract_block("conformation", 0, blob);
        BlobData B;
        std::string f_nodes, f_top, f_surf, f_mat, f_stokes, f_pin, motion;
        bool set_ssint = false;
        for (const auto &l : conf) {
            std::string lv, rv;
            parse_tag(l, lv, rv);
            if (lv == "motion_state") motion = rv;
            else if (lv == "nodes") f_nodes = join(dir, rv);
            else if (lv == "topology") f_top = join(dir, rv);
            else if (lv == "surface") f_surf = join(dir, rv);
            else if (lv == "material") f_mat = join(dir, rv);
            else if (lv == "stokes") f_stokes = join(dir, rv);
            else if (lv == "vdw" || lv == "ssint") set_ssint = true;
            else if (lv == "pin") f_pin = join(dir, rv);
        }
        std::string solver = "CG_nomass";
        bool set_scale = false;
        // blob-level tags, skipping the conformation blocks (World.cpp:2592-2690)
        int depth = 0;
        for (const auto &l : blob) {
            const std::string t = tag_name(l);
            if (t == "conformation" || t == "kinetics") ++depth;
            if (t == "/conformation" || t == "/kinetics") {
                --depth;
                continue;
            }
            if (depth > 0) continue;
            std::string lv, rv;
            if (!parse_tag(l, lv, rv)) continue;
            if (lv == "solver") solver = rv;
            else if (lv == "scale") {
                B.scale = atof(rv.c_str()) / meso::length;
                set_scale = true;
            } else if (lv == "centroid" || lv == "centroid_pos") {
                parse_vec3(rv, B.centroid);
                B.set_centroid = true;
            }
        }
        if (motion.empty())
            fail("In blob %d, conformation 0: must have a motion state set.\nAccepted states: DYNAMIC, STATIC, FROZEN.",
                 b);
        if (motion != "DYNAMIC") fail("In blob %d: only DYNAMIC blobs are handled by this build.", b);
        if (f_nodes.empty() || f_surf.empty() || !set_ssint)
            fail("In blob %d, conformation 0:\nFor any blob conformation, 'nodes', 'surface' and 'ssint' must be set.",
                 b);
        if (f_top.empty() || f_mat.empty() || f_stokes.empty())
            fail("In blob %d, conformation 0:\nFor a DYNAMIC blob conformation, 'topology', 'material', and 'stokes' "
                 "must be set.",
                 b);
        if (!set_scale) fail("Blob %d, scale not set.", b);
        // World.cpp:2600-2625
        if (solver != "CG_nomass" && solver != "masslumped" && solver != "CG" && solver != "direct")
            fail("In blob %d, unrecognised solver type.\nRecognised solvers:CG, CG_nomass, direct, masslumped.", b);
        if (solver == "direct")
            fail("Blob %d: solver 'direct' is not part of this build (CG_nomass, CG and masslumped are).", b);
        B.solver = solver == "masslumped" ? FFEA_GPU_SOLVER_MASSLUMPED
                   : solver == "CG"       ? FFEA_GPU_SOLVER_ITERATIVE
                                          : FFEA_GPU_SOLVER_NOMASS_CG;
        // a HARD wall zeroes outward velocities of surface nodes before each step
        // (Blob::enforce_box_boundaries, Blob.cpp:1800-1845): the mass-less solve overwrites
        // them, an inertial blob would keep them; that is not built
        if (P.hard_wall && B.solver != FFEA_GPU_SOLVER_NOMASS_CG)
            fail("Blob %d: HARD walls with an inertial solver are not part of this build.", b);
        load_nodes(B, f_nodes);
        load_topology(B, f_top);
        load_surface(B, f_surf);
        load_material(B, f_mat);
        load_stokes(B, f_stokes);
        if (!ctf_file.empty()) load_ctforces(B, b, ctf_file);
        if (!f_pin.empty()) load_pins(B, f_pin);
        // Blob::init linearises every element (Blob.cpp:203-205)
        static const int ends[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        for (int e = 0; e < B.ne; ++e) {
            const int *n = &B.top[(size_t)10 * e];
            for (int k = 0; k < 6; ++k)
                for (int c = 0; c < 3; ++c)
                    B.pos[3 * n[4 + k] + c] = .5 * (B.pos[3 * n[ends[k][0]] + c] + B.pos[3 * n[ends[k][1]] + c]);
        }
        // Blob::position (Blob.cpp:613-650) for a given centroid
        if (B.set_centroid) {
            double c[3] = {0, 0, 0};
            for (int i = 0; i < B.nt; ++i)
                for (int k = 0; k < 3; ++k) c[k] += B.pos[3 * i + k];
            for (int k = 0; k < 3; ++k) c[k] *= (1.0 / B.nt);
            for (int i = 0; i < B.nt; ++i)
                for (int k = 0; k < 3; ++k) B.pos[3 * i + k] += B.centroid[k] * B.scale - c[k];
        }
        S.blobs.push_back(std::move(B));
    }
    if (p.calc_springs == 1) {  // World.cpp:2891-2896, after the blobs
        if (!springs_given) fail("Failed to open springs file: '' (no <springs> block in <interactions>)");
        load_springs(S, springs_file);
    }
    // World.cpp:442-474: box from the system size when no grid is set, then the system
    // centroid moved to the box centre
    double wc[3] = {0, 0, 0};
    long long ntot = 0;
    double mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = std::numeric_limits<double>::max();
        mx[k] = std::numeric_limits<double>::min();  // as World::get_system_dimensions initialises it
    }
    for (const auto &B : S.blobs) {
        double c[3] = {0, 0, 0};
        for (int i = 0; i < B.nt; ++i)
            for (int k = 0; k < 3; ++k) {
                c[k] += B.pos[3 * i + k];
                mn[k] = std::min(mn[k], B.pos[3 * i + k]);
                mx[k] = std::max(mx[k], B.pos[3 * i + k]);
            }
        for (int k = 0; k < 3; ++k) wc[k] += (c[k] / B.nt) * B.nt;
        ntot += B.nt;
    }
    for (int k = 0; k < 3; ++k) wc[k] /= ntot;
    int esN[3] = {P.es_N[0], P.es_N[1], P.es_N[2]};
    if (esN[0] < 1 || esN[1] < 1 || esN[2] < 1)
        for (int k = 0; k < 3; ++k) esN[k] = 2 * (int)ceil((mx[k] - mn[k]) / P.ssint_cutoff);
    if (p.move_into_box == 1)
        for (auto &B : S.blobs)
            for (int i = 0; i < B.nt; ++i)
                for (int k = 0; k < 3; ++k) B.pos[3 * i + k] += P.ssint_cutoff * esN[k] / 2.0 - wc[k];
    // World.cpp:269-276: 6 seeds from glibc rand; the single-thread stream starts there
    srand((unsigned)p.rng_seed);
    const uint32_t two_max[2] = {4294967087u, 4294944443u};
    for (int i = 0; i < 6; ++i) {
        const uint32_t a = (uint32_t)rand(), b2 = (uint32_t)rand();
        S.rng_state[i] = (a + b2) % two_max[i / 3];
    }
}

}  // namespace

bool ffea_internal_script_files(const ffea_script *s, ScriptFiles *out) {
    if (!s || !out) return false;
    const std::string dir = dirname_of(s->path);
    std::string base = s->path.substr(s->path.find_last_of('/') == std::string::npos ? 0 : s->path.find_last_of('/') + 1);
    const size_t dot = base.find_last_of('.');
    const std::string stem = dot == std::string::npos ? base : base.substr(0, dot);
    const Params &P = s->P;
    out->script = s->path;
    out->trajectory = P.traj_out.empty() ? dir + "/" + stem + "/.ftj" : join(dir, P.traj_out);
    out->measurement = P.meas_out.empty() ? dir + "/" + stem + "/.fm" : join(dir, P.meas_out);
    // the detailed measurement file (SimulationParams.cpp:690-709): det_measurement_out_fname,
    // else the measurement file's path with its extension replaced by .fdm; with neither
    // given the reference's name stays "\n" and no detailed file is written here
    if (!P.det_out.empty()) {
        out->detailed = join(dir, P.det_out);
    } else if (!P.meas_out.empty()) {
        const std::string &m = out->measurement;
        const size_t sl = m.find_last_of('/'), d = m.find_last_of('.');
        const bool ext = d != std::string::npos && (sl == std::string::npos ? d > 0 : d > sl + 1);
        out->detailed = (ext ? m.substr(0, d) : m) + ".fdm";
    } else {
        out->detailed.clear();
    }
    out->checkpoint_in = P.ckpt_in.empty() ? std::string() : join(dir, P.ckpt_in);
    if (!P.ckpt_out.empty()) out->checkpoint_out = join(dir, P.ckpt_out);
    else out->checkpoint_out = (dot == std::string::npos ? s->path : s->path.substr(0, s->path.size() - (base.size() - dot))) + ".fcp";
    return true;
}

extern "C" {

int ffea_script_load(const char *path, ffea_script **out) {
    if (!path || !out) {
        ffea_internal_set_error("ffea_script_load: invalid argument");
        return FFEA_GPU_ERR_INVALID;
    }
    *out = nullptr;
    std::unique_ptr<ffea_script> s(new ffea_script());
    s->path = path;
    try {
        load_script(*s);
    } catch (const Error &e) {
        ffea_internal_set_error(e.msg);
        return FFEA_GPU_ERR_INVALID;
    } catch (const std::bad_alloc &) {
        ffea_internal_set_error("host allocation failed");
        return FFEA_GPU_ERR_INVALID;
    }
    *out = s.release();
    return FFEA_GPU_OK;
}

void ffea_script_free(ffea_script *s) { delete s; }

int ffea_script_get_params(const ffea_script *s, ffea_script_params *out) {
    if (!s || !out) return FFEA_GPU_ERR_INVALID;
    *out = s->P.p;
    return FFEA_GPU_OK;
}

int ffea_script_blob_sizes(const ffea_script *s, int b, int *nn, int *ne, int *np, int *nc, int *nf) {
    if (!s || b < 0 || b >= (int)s->blobs.size()) return FFEA_GPU_ERR_INVALID;
    const BlobData &B = s->blobs[b];
    if (nn) *nn = B.nt;
    if (ne) *ne = B.ne;
    if (np) *np = (int)B.pinned.size();
    if (nc) *nc = (int)B.ctf_node.size();
    if (nf) *nf = B.num_surface_faces;
    return FFEA_GPU_OK;
}

int ffea_script_blob_arrays(const ffea_script *s, int b, double *pos, int *top, double *A, double *Bv, double *G,
                            double *E, double *rho, int *pinned, int *ctf_node, double *ctf_force,
                            double *stokes_radius) {
    if (!s || b < 0 || b >= (int)s->blobs.size()) return FFEA_GPU_ERR_INVALID;
    const BlobData &B = s->blobs[b];
    auto cp = [](auto *dst, const auto &src) {
        if (dst) std::copy(src.begin(), src.end(), dst);
    };
    cp(pos, B.pos);
    cp(top, B.top);
    cp(A, B.A);
    cp(Bv, B.B);
    cp(G, B.G);
    cp(E, B.E);
    cp(rho, B.rho);
    cp(pinned, B.pinned);
    cp(ctf_node, B.ctf_node);
    cp(ctf_force, B.ctf_force);
    cp(stokes_radius, B.stokes_radius);
    return FFEA_GPU_OK;
}

int ffea_script_blob_surface_forces(const ffea_script *s, int b, int *num_faces, int *num_sets, int *num_entries,
                                    int *face_nodes, double *set_force, int *set_ptr, int *set_faces) {
    if (!s || b < 0 || b >= (int)s->blobs.size()) return FFEA_GPU_ERR_INVALID;
    const BlobData &B = s->blobs[b];
    if (num_faces) *num_faces = B.num_surface_faces;
    if (num_sets) *num_sets = (int)B.sf_ptr.size() - 1;
    if (num_entries) *num_entries = (int)B.sf_faces.size();
    auto cp = [](auto *dst, const auto &src) {
        if (dst) std::copy(src.begin(), src.end(), dst);
    };
    cp(face_nodes, B.surf_nodes);
    cp(set_force, B.sf_force);
    cp(set_ptr, B.sf_ptr);
    cp(set_faces, B.sf_faces);
    return FFEA_GPU_OK;
}

int ffea_script_blob_rotational_forces(const ffea_script *s, int b, int *num, int *node, double *force, double *axis,
                                       char *type) {
    if (!s || b < 0 || b >= (int)s->blobs.size()) return FFEA_GPU_ERR_INVALID;
    const BlobData &B = s->blobs[b];
    if (num) *num = (int)B.rf_node.size();
    auto cp = [](auto *dst, const auto &src) {
        if (dst) std::copy(src.begin(), src.end(), dst);
    };
    cp(node, B.rf_node);
    cp(force, B.rf_force);
    cp(axis, B.rf_axis);
    cp(type, B.rf_type);
    return FFEA_GPU_OK;
}

int ffea_script_blob_solver(const ffea_script *s, int b, int *solver) {
    if (!s || !solver || b < 0 || b >= (int)s->blobs.size()) return FFEA_GPU_ERR_INVALID;
    *solver = s->blobs[b].solver;
    return FFEA_GPU_OK;
}

int ffea_script_springs(const ffea_script *s, int *num, double *k, double *l, int *blob, int *conf, int *node,
                        int *active) {
    if (!s) return FFEA_GPU_ERR_INVALID;
    if (num) *num = (int)s->springs.size();
    for (size_t i = 0; i < s->springs.size(); ++i) {
        const SpringData &sp = s->springs[i];
        if (k) k[i] = sp.k;
        if (l) l[i] = sp.l;
        for (int j = 0; j < 2; ++j) {
            if (blob) blob[2 * i + j] = sp.blob[j];
            if (conf) conf[2 * i + j] = sp.conf[j];
            if (node) node[2 * i + j] = sp.node[j];
        }
        if (active) active[i] = sp.active ? 1 : 0;
    }
    return FFEA_GPU_OK;
}

int ffea_script_rng_state(const ffea_script *s, uint32_t state[6]) {
    if (!s || !state) return FFEA_GPU_ERR_INVALID;
    std::copy(s->rng_state, s->rng_state + 6, state);
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_create_from_script(const ffea_script *s, int b, int precision, int device, ffea_gpu_blob **out) {
    if (!s || !out || b < 0 || b >= (int)s->blobs.size()) {
        ffea_internal_set_error("ffea_gpu_blob_create_from_script: invalid argument");
        return FFEA_GPU_ERR_INVALID;
    }
    const BlobData &B = s->blobs[b];
    const ffea_script_params &p = s->P.p;
    if (p.calc_ssint || p.calc_steric || p.calc_es || p.calc_preComp || p.calc_kinetics) {
        ffea_internal_set_error("surface interactions, electrostatics, pre-computed potentials and kinetics are not "
                                "part of this build; set calc_ssint, calc_steric, calc_es, calc_preComp and "
                                "calc_kinetics to 0");
        return FFEA_GPU_ERR_INVALID;
    }
    ffea_gpu_params gp{};
    gp.dt = p.dt;
    gp.kT = p.kT;
    gp.epsilon2 = p.epsilon2;
    gp.max_iterations_cg = p.max_iterations_cg;
    gp.calc_noise = p.calc_noise;
    gp.calc_stokes = p.calc_stokes;
    gp.precision = precision;
    gp.device = device;
    // Blob::init (Blob.cpp:264-268): drag = 6 pi stokes_visc stokes_radius per node
    const double pi = 3.141592653589793238462643383279502884;
    std::vector<double> drag;
    if (p.calc_stokes) {
        drag.resize(B.nt);
        for (int i = 0; i < B.nt; ++i) drag[i] = 6.0 * pi * p.stokes_visc * B.stokes_radius[i];
    }
    ffea_gpu_surface_forces sf{};
    const int nsets = (int)B.sf_ptr.size() - 1;
    if (nsets > 0) {
        sf.num_faces = B.num_surface_faces;
        sf.face_nodes = B.surf_nodes.data();
        sf.num_sets = nsets;
        sf.set_force = B.sf_force.data();
        sf.set_ptr = B.sf_ptr.data();
        sf.set_faces = B.sf_faces.data();
    }
    ffea_gpu_rotational_forces rf{(int)B.rf_node.size(), B.rf_node.data(), B.rf_force.data(), B.rf_axis.data(),
                                  B.rf_type.data()};
    std::vector<int> ext;  // the nodes of this blob that active springs pull (World::apply_springs)
    for (const SpringData &sp : s->springs)
        for (int j = 0; j < 2; ++j)
            if (sp.active && sp.blob[j] == b) ext.push_back(sp.node[j]);
    int rc = ffea_gpu_blob_create_ex(&gp, B.nt, B.ne, B.pos.data(), B.top.data(), B.A.data(), B.B.data(), B.G.data(),
                                     B.E.data(), (int)B.pinned.size(), B.pinned.empty() ? nullptr : B.pinned.data(),
                                     (int)B.ctf_node.size(), B.ctf_node.empty() ? nullptr : B.ctf_node.data(),
                                     B.ctf_force.empty() ? nullptr : B.ctf_force.data(),
                                     p.calc_stokes ? drag.data() : nullptr, s->rng_state, nsets > 0 ? &sf : nullptr,
                                     B.rf_node.empty() ? nullptr : &rf, (int)ext.size(),
                                     ext.empty() ? nullptr : ext.data(), out);
    if (rc == FFEA_GPU_OK && B.solver != FFEA_GPU_SOLVER_NOMASS_CG) {
        rc = ffea_gpu_blob_set_solver(*out, B.solver, B.rho.data());
        if (rc != FFEA_GPU_OK) {
            ffea_internal_set_error(ffea_gpu_blob_last_error(*out));
            ffea_gpu_blob_destroy(*out);
            *out = nullptr;
            return rc;
        }
    }
    if (rc != FFEA_GPU_OK || s->blobs.size() == 1) return rc;
    // several blobs draw from the one stream in turn (World.cpp:2031-2058): all element
    // draws first, blob after blob, then (Stokes noise) all node draws, blob after blob
    const bool node_draws = p.calc_stokes && p.calc_noise;
    long long before = 0, total = 0, elem_all = 0, nodes_before = 0;
    for (size_t i = 0; i < s->blobs.size(); ++i) {
        if ((int)i < b) {
            before += 7ll * s->blobs[i].ne;
            nodes_before += 3ll * s->blobs[i].nt;
        }
        elem_all += 7ll * s->blobs[i].ne;
        total += 7ll * s->blobs[i].ne + (node_draws ? 3ll * s->blobs[i].nt : 0ll);
    }
    if (!p.calc_noise) return rc;
    uint32_t st[6];
    ffea_script_rng_state(s, st);
    rc = ffea_gpu_advance_rng_state(st, before, st);
    if (rc == FFEA_GPU_OK) rc = ffea_gpu_blob_set_rng_stream(*out, st, total);
    if (rc == FFEA_GPU_OK && node_draws) rc = ffea_gpu_batch_set_node_draw_offset(*out, 0, elem_all + nodes_before - before);
    if (rc != FFEA_GPU_OK) {
        ffea_gpu_blob_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

}  // extern "C"
