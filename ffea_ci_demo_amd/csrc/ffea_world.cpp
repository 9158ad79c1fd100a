// The World loop of an FFEA script over GPU blob handles: trajectory (.ftj), global
// measurement (.fm), detailed measurement (.fdm) and checkpoint (.fcp) files in the
// reference's formats, and restart from them. One handle per blob; each blob keeps its own positions on the GPU and the
// host only sees them at output steps.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "../../include/ffea_gpu.h"
#include "../../include/ffea_script.h"
#include "../../include/ffea_world.h"
#include "script_internal.h"

void ffea_internal_set_error(const std::string &msg);  // ffea_blob.hip

namespace {

// mesoDimensions (dimensions.h): length, Energy and time in SI
constexpr double kLength = 1.7e-10;
constexpr double kEnergy = 4.141945559999999e-21;
constexpr double kTime = 3.7302915560886126e-13;
constexpr double kPressure = 8.4305832688784826e8;
constexpr double kVelocity = 4.5572845297447225e2;
constexpr double kForce = 2.4364705882352941e-11;

// one rotational constant force on a node at x (Blob::apply_ctforces CTF 2,
// Blob.cpp:1693-1714; intersectingPointToLine, mat_vec_fns_II.h:755-777), for the force
// column of an inertial blob's trajectory frame
void rotational_force(const double *x, const double *ax, double F, bool torque, double *f) {
    const double *a = ax, *d = ax + 3;
    auto cross = [](const double *u, const double *v, double *w) {
        w[0] = u[1] * v[2] - v[1] * u[2];
        w[1] = -u[0] * v[2] + v[0] * u[2];
        w[2] = u[0] * v[1] - v[0] * u[1];
    };
    double pa[3], v2[3], cp[3], w[3], r[3], g[3];
    for (int c = 0; c < 3; ++c) pa[c] = x[c] - a[c];
    cross(pa, d, v2);
    double m = 0.0;
    for (int c = 0; c < 3; ++c) m += v2[c] * v2[c];
    if (std::sqrt(m) < 3.0 * 2.220446049250313e-16) {
        for (int c = 0; c < 3; ++c) r[c] = x[c] - x[c];
    } else {
        cross(d, v2, cp);
        cross(d, cp, w);
        double det = pa[0] * (d[1] * w[2] - d[2] * w[1]);
        det += pa[1] * (d[2] * w[0] - d[0] * w[2]);
        det += pa[2] * (d[0] * w[1] - d[1] * w[0]);
        const double cc = det / (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        for (int c = 0; c < 3; ++c) r[c] = (x[c] + cc * cp[c]) - x[c];
    }
    cross(r, d, g);
    double s = F;
    if (torque) {
        double rr = 0.0;
        for (int c = 0; c < 3; ++c) rr += r[c] * r[c];
        s = F / std::sqrt(rr);
    }
    for (int c = 0; c < 3; ++c) f[c] = s * g[c];
}

struct WorldError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg) { throw WorldError{code, msg}; }

bool file_exists(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

// SimulationParams::checkFileName (SimulationParams.cpp:779-802): an existing output file
// is renamed to __<name>__bckp.N in its directory, N the first number not yet taken
void move_aside(const std::string &path) {
    if (!file_exists(path)) return;
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    const std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    const std::string base = dir + "__" + name + "__bckp.";
    int n = 1;
    while (file_exists(base + std::to_string(n))) ++n;
    if (rename(path.c_str(), (base + std::to_string(n)).c_str()) != 0)
        fail(FFEA_GPU_ERR_IO, "could not move " + path + " aside: " + strerror(errno));
    fprintf(stderr, "FFEA: moving %s to %s\n", path.c_str(), (base + std::to_string(n)).c_str());
}

std::string read_whole(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) fail(FFEA_GPU_ERR_IO, "failed to open " + path);
    std::string s;
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
    fclose(f);
    return s;
}

void truncate_to(const std::string &path, long long size) {
    if (truncate(path.c_str(), (off_t)size) != 0)
        fail(FFEA_GPU_ERR_IO, "error when trying to truncate " + path + ": " + strerror(errno));
}

std::vector<std::string> split_spaces(const std::string &line) {  // boost::split on single spaces
    std::vector<std::string> out(1);
    for (char c : line) {
        if (c == ' ') out.emplace_back();
        else out.back() += c;
    }
    return out;
}

}  // namespace

struct ffea_world {
    ScriptFiles files;
    ffea_script_params p{};
    int precision = FFEA_GPU_FP64;
    std::vector<ffea_gpu_blob *> blobs;
    std::vector<int> nt;
    std::vector<std::vector<double>> pos0;  // positions at set-up, the RMSD origin (World.cpp:505-513)
    std::vector<long long> draws_before;   // noise draws of the blobs before this one in a step
    std::vector<std::string> old_seeds;     // extra seed lines of an earlier multi-threaded run
    FILE *traj = nullptr, *meas = nullptr, *ckpt = nullptr;
    FILE *det = nullptr;   // the detailed measurement file (.fdm), unless turned off (the reference's -d)
    bool detailed = true;
    int frames_to_delete = 0;  // restart: complete frames dropped before resuming (the reference's -l)
    long long step = 0, step_initial = 0;
    std::vector<double> host;
    std::string err;
    // springs (World::load_springs, World.cpp:3543-3643), file order
    struct Spring {
        double k = 0, l = 0;
        int blob[2] = {0, 0}, node[2] = {0, 0};
        bool active = false;
    };
    std::vector<Spring> springs;
    // inertial blobs (<solver = masslumped>): the trajectory carries velocities and the
    // node forces of the print, the measurements the kinetic energy
    std::vector<char> inertial;
    bool mass_in_system = false;
    struct NodeForces {  // the constant forces of a blob, list order
        std::vector<int> lin_node, rot_node;
        std::vector<double> lin_force, rot_force, rot_axis;
        std::vector<char> rot_type;
    };
    std::vector<NodeForces> ctf;
    std::vector<std::vector<int>> spring_node;       // this step's spring forces per blob, spring order
    std::vector<std::vector<double>> spring_force;

    ~ffea_world() {
        for (FILE *f : {traj, meas, ckpt, det})
            if (f) fclose(f);
        for (ffea_gpu_blob *b : blobs) ffea_gpu_blob_destroy(b);
    }

    void check_rc(int rc, int blob, const char *what) {
        if (rc == FFEA_GPU_OK) return;
        const char *m = blob >= 0 ? ffea_gpu_blob_last_error(blobs[blob]) : ffea_gpu_last_error();
        fail(rc, std::string(what) + (blob >= 0 ? " (blob " + std::to_string(blob) + ")" : std::string()) + ": " +
                     (m ? m : ""));
    }

    // positions [n][3] of the given nodes of blob b
    void node_positions(int b, const std::vector<int> &nodes, std::vector<double> &x) {
        x.resize(3 * nodes.size());
        if (!nodes.empty())
            check_rc(ffea_gpu_blob_get_node_positions(blobs[b], (int)nodes.size(), nodes.data(), x.data()), b,
                     "reading spring node positions");
    }

    // World::apply_springs (World.cpp:4063-4108): each active spring pulls its two nodes with
    // k (|x1 - x0| - l) along (x1 - x0)/|x1 - x0| (normalize and magnitude,
    // mat_vec_fns_II.h:188-198, 326-332), in spring order; the forces reach the blobs through
    // Blob::add_force_to_node before their update_internal_forces
    void apply_springs() {
        if (springs.empty()) return;
        const int nb = (int)blobs.size();
        std::vector<std::vector<int>> want(nb);
        for (const Spring &sp : springs)
            if (sp.active)
                for (int j = 0; j < 2; ++j) want[sp.blob[j]].push_back(sp.node[j]);
        std::vector<std::vector<double>> x(nb);
        for (int b = 0; b < nb; ++b) node_positions(b, want[b], x[b]);
        std::vector<std::vector<int>> fn(nb);
        std::vector<std::vector<double>> ff(nb);
        std::vector<size_t> used(nb, 0);
        for (const Spring &sp : springs) {
            if (!sp.active) continue;
            // the ends were listed in spring order, end 0 first
            const double *n0 = &x[sp.blob[0]][3 * used[sp.blob[0]]++];
            const double *n1 = &x[sp.blob[1]][3 * used[sp.blob[1]]++];
            double sep[3] = {n1[0] - n0[0], n1[1] - n0[1], n1[2] - n0[2]}, sep_norm[3];
            double sum = 0;
            for (int c = 0; c < 3; ++c) sum += sep[c] * sep[c];
            // a zero separation: normalize throws, and World::apply_springs's catch (int) lets it through
            if (sum == 0) fail(FFEA_GPU_ERR_INVALID, "Unable to normalize a vector with length zero");
            sum = std::sqrt(sum);
            for (int c = 0; c < 3; ++c) sep_norm[c] = sep[c] / sum;
            double mag = 0.0;
            for (int c = 0; c < 3; ++c) mag += sep[c] * sep[c];
            const double force_mag = sp.k * (std::sqrt(mag) - sp.l);
            fn[sp.blob[0]].push_back(sp.node[0]);
            for (int c = 0; c < 3; ++c) ff[sp.blob[0]].push_back(force_mag * sep_norm[c]);
            fn[sp.blob[1]].push_back(sp.node[1]);
            for (int c = 0; c < 3; ++c) ff[sp.blob[1]].push_back(-1 * force_mag * sep_norm[c]);
        }
        for (int b = 0; b < nb; ++b)
            if (!fn[b].empty())
                check_rc(ffea_gpu_blob_add_node_forces(blobs[b], (int)fn[b].size(), fn[b].data(), ff[b].data()), b,
                         "applying springs");
        spring_node = std::move(fn);
        spring_force = std::move(ff);
    }

    // Blob::force at the print (World.cpp:2045-2049), an inertial blob's force column: zeroed
    // (World.cpp:1862), the springs' pulls (:1910), then the linear and rotational constant
    // forces (Blob::apply_ctforces, Blob.cpp:1679-1719) at the frame's positions
    void print_forces(int b, const std::vector<double> &x, std::vector<double> &f) {
        f.assign(3 * (size_t)nt[b], 0.0);
        if (b < (int)spring_node.size())
            for (size_t i = 0; i < spring_node[b].size(); ++i)
                for (int c = 0; c < 3; ++c) f[3 * (size_t)spring_node[b][i] + c] += spring_force[b][3 * i + c];
        const NodeForces &F = ctf[b];
        for (size_t i = 0; i < F.lin_node.size(); ++i)
            for (int c = 0; c < 3; ++c) f[3 * (size_t)F.lin_node[i] + c] += F.lin_force[3 * i + c];
        for (size_t i = 0; i < F.rot_node.size(); ++i) {
            double g[3];
            rotational_force(&x[3 * (size_t)F.rot_node[i]], &F.rot_axis[6 * i], F.rot_force[i],
                             F.rot_type[2 * i + 1] == 't', g);
            for (int c = 0; c < 3; ++c) f[3 * (size_t)F.rot_node[i] + c] += g[c];
        }
    }

    // the spring part of World::make_measurements (World.cpp:4632-4652): every spring, active
    // or not, adds 0.5 k (|x0 - x1| - l)^2 to its (blob 0, blob 1) cell [nb][nb]
    std::vector<double> spring_cells() {
        const int nb = (int)blobs.size();
        std::vector<std::vector<int>> want(nb);
        for (const Spring &sp : springs)
            for (int j = 0; j < 2; ++j) want[sp.blob[j]].push_back(sp.node[j]);
        std::vector<std::vector<double>> x(nb);
        for (int b = 0; b < nb; ++b) node_positions(b, want[b], x[b]);
        std::vector<double> cell((size_t)nb * nb, 0.0);
        std::vector<size_t> used(nb, 0);
        for (const Spring &sp : springs) {
            const double *a = &x[sp.blob[0]][3 * used[sp.blob[0]]++];
            const double *b = &x[sp.blob[1]][3 * used[sp.blob[1]]++];
            double m = 0.0;
            for (int c = 0; c < 3; ++c) m += (a[c] - b[c]) * (a[c] - b[c]);
            m = std::sqrt(m);
            cell[(size_t)sp.blob[0] * nb + sp.blob[1]] += 0.5 * sp.k * (m - sp.l) * (m - sp.l);
        }
        return cell;
    }

    // the whole field: the cells summed row by row (World::get_spring_field_energy(-1, -1),
    // World.cpp:4110-4123)
    static double spring_total(const std::vector<double> &cell) {
        double e = 0.0;
        for (double v : cell) e += v;
        return e;
    }

    // a blob with a spring at either end (World::load_springs, World.cpp:3596-3597)
    bool has_springs(int b) const {
        for (const Spring &sp : springs)
            if (sp.blob[0] == b || sp.blob[1] == b) return true;
        return false;
    }

    // World::init headers (World.cpp:527-603)
    void write_headers(const char *script_path) {
        fprintf(traj, "FFEA_trajectory_file\n\nInitialisation:\nNumber of Blobs %d\nNumber of Conformations",
                p.num_blobs);
        for (int i = 0; i < p.num_blobs; ++i) fprintf(traj, " %d", 1);
        fprintf(traj, "\n");
        for (int i = 0; i < p.num_blobs; ++i) fprintf(traj, "Blob %d:\tConformation %d Nodes %d\t\n", i, 0, nt[i]);
        fprintf(traj, "\n*\n");
        fflush(traj);

        fprintf(meas, "FFEA Global Measurement File\n\nSimulation Details:\n");
        fprintf(meas, "\tGPU blob path: libffea_gpu (HIP, gfx950), %s\n",
                precision == FFEA_GPU_MIXED ? "mixed FP32/FP64" : "FP64");
        const time_t now = time(nullptr);
        const tm *t = localtime(&now);
        fprintf(meas, "\tSimulation Began on %d/%d/%d at %d:%d:%d\n", t->tm_mday, 1 + t->tm_mon, 1900 + t->tm_year,
                t->tm_hour, t->tm_min, t->tm_sec);
        fprintf(meas, "\tScript Filename = %s\n", script_path);
        fprintf(meas, "\tSimulation Type = %s\n\n", "Full");
        // SimulationParams::write_to_file (SimulationParams.cpp:1075-1233), the keys of this path
        fprintf(meas, "Parameters:\n\tSystem parameters:\n");
        fprintf(meas, "\trestart = %d\n", p.restart);
        fprintf(meas, "\tdt = %e\n", p.dt * kTime);
        fprintf(meas, "\tnum_steps = %lld\n", p.num_steps);
        fprintf(meas, "\tcheck = %d\n", p.check);
        fprintf(meas, "\trng_seed = %d\n", p.rng_seed);
        fprintf(meas, "\tkT = %e\n", p.kT * kEnergy);
        fprintf(meas, "\tmax_iterations_cg = %d\n", p.max_iterations_cg);
        fprintf(meas, "\tepsilon = %e\n", std::sqrt(p.epsilon2));
        if (p.calc_stokes == 1) fprintf(meas, "\tstokes_visc = %e\n", p.stokes_visc * kPressure * kTime);
        fprintf(meas, "\tnum_blobs = %d\n", p.num_blobs);
        fprintf(meas, "\tnum_rods = %d\n", 0);
        fprintf(meas, "\n\tCalculations enabled:\n");
        fprintf(meas, "\tcalc_noise = %d\n\tcalc_stokes = %d\n", p.calc_noise, p.calc_stokes);
        fprintf(meas, "\tcalc_ssint = %d\n\tcalc_steric = %d\n\tcalc_preComp = %d\n\tcalc_springs = %d\n", 0, 0, 0,
                p.calc_springs);
        fprintf(meas, "\tcalc_ctforces = %d\n\tcalc_kinetics = %d\n\tcalc_es = %d\n", p.calc_ctforces, 0, 0);
        fprintf(meas, "\n\tSimulation box configuration:\n\tmove_into_box = %d\n", p.move_into_box);
        fprintf(meas, "\n\n");
        fprintf(meas, "Measurements:\n");
        fprintf(meas, "%-14s", "Time");
        if (mass_in_system) fprintf(meas, "%-14s", "KineticEnergy");
        fprintf(meas, "%-14s", "StrainEnergy");
        fprintf(meas, "%-14s%-14s%-14s%-14s", "Centroid[0]", "Centroid[1]", "Centroid[2]", "RMSD");
        if (p.calc_springs != 0) fprintf(meas, "%-14s", "SpringEnergy");
        fprintf(meas, "\n");
        fflush(meas);
        if (det) write_detailed_header();
    }

    // the detailed measurement file's header (World.cpp:606-646): per blob its kinetic energy
    // (blobs with mass), strain energy, centroid and RMSD; per blob pair (i <= j) its spring
    // energy where both blobs carry springs, the pair block only with springs on (surface
    // interactions and PreComp are off on this path)
    void write_detailed_header() {
        fprintf(det, "FFEA Detailed Measurement File\n\nMeasurements:\n");
        fprintf(det, "%-14s", "Time");
        for (int i = 0; i < p.num_blobs; ++i) {
            fprintf(det, "| B%d ", i);
            if (inertial[i]) fprintf(det, "%-14s", "KineticEnergy");
            fprintf(det, "%-14s", "StrainEnergy");
            fprintf(det, "%-14s%-14s%-14s%-14s", "Centroid[0]", "Centroid[1]", "Centroid[2]", "RMSD");
        }
        if (p.calc_springs == 1)
            for (int i = 0; i < p.num_blobs; ++i)
                for (int j = i; j < p.num_blobs; ++j) {
                    fprintf(det, "| B%dB%d ", i, j);
                    if (has_springs(i) && has_springs(j)) fprintf(det, "%-14s", "SpringEnergy");
                }
        fprintf(det, "\n");
        fflush(det);
    }

    // World::print_checkpoints (World.cpp:4546-4596): the noise stream's state before this
    // step's forces; the blobs share it in turn, blob 0 first, so blob 0's state is the stream's
    void print_checkpoint() {
        uint32_t s[6];
        check_rc(ffea_gpu_blob_get_rng_state(blobs[0], s), 0, "reading the noise stream state");
        rewind(ckpt);
        fprintf(ckpt, "RNGStreams dedicated to the thermal stress: %d\n", 1 + (int)old_seeds.size());
        fprintf(ckpt, "%u %u %u %u %u %u\n", s[0], s[1], s[2], s[3], s[4], s[5]);
        for (const std::string &l : old_seeds) fprintf(ckpt, "%s\n", l.c_str());
        fflush(ckpt);
    }

    // trajectory frame (World.cpp:4461-4488, Blob::write_nodes_to_file Blob.cpp:855-889)
    // and global measurements (World::make_measurements, write_measurements_to_file,
    // World.cpp:4598-4692; Blob::make_measurements Blob.cpp:1066-1202)
    void print_outputs(long long s) {
        double strain = 0.0, kinetic = 0.0, cog[3] = {0, 0, 0}, rmsd2 = 0.0;
        long long nodes = 0;
        // per blob (Blob::make_measurements): kinetic and strain energy, centroid, RMSD
        std::vector<double> b_ke(p.num_blobs, 0.0), b_strain(p.num_blobs, 0.0), b_cog(3 * (size_t)p.num_blobs, 0.0),
            b_rmsd(p.num_blobs, 0.0);
        std::vector<double> vel, frc;
        for (int b = 0; b < p.num_blobs; ++b) {
            host.resize(3 * (size_t)nt[b]);
            check_rc(ffea_gpu_blob_get_positions(blobs[b], host.data()), b, "reading positions");
            if (inertial[b]) {
                vel.resize(3 * (size_t)nt[b]);
                check_rc(ffea_gpu_blob_get_velocities(blobs[b], vel.data()), b, "reading velocities");
                print_forces(b, host, frc);
                double ke = 0.0;
                check_rc(ffea_gpu_blob_kinetic_energy(blobs[b], &ke), b, "kinetic energy");
                kinetic += ke;
                b_ke[b] = ke;
            }
            fprintf(traj, "Blob %d, Conformation %d, step %d\n", b, 0, (int)s);
            fprintf(traj, "DYNAMIC\n");
            double c[3] = {0, 0, 0}, r = 0.0;
            for (int i = 0; i < nt[b]; ++i) {
                const double *x = &host[3 * (size_t)i];
                if (inertial[b]) {  // Blob::write_nodes_to_file (Blob.cpp:866-873)
                    const double *v = &vel[3 * (size_t)i], *f = &frc[3 * (size_t)i];
                    fprintf(traj, "%e %e %e %e %e %e %e %e %e %e\n", x[0] * kLength, x[1] * kLength, x[2] * kLength,
                            v[0] * kVelocity, v[1] * kVelocity, v[2] * kVelocity, 0., f[0] * kForce, f[1] * kForce,
                            f[2] * kForce);
                } else {
                    fprintf(traj, "%e %e %e %e %e %e %e %e %e %e\n", x[0] * kLength, x[1] * kLength, x[2] * kLength, 0.,
                            0., 0., 0., 0., 0., 0.);
                }
                for (int k = 0; k < 3; ++k) {
                    c[k] += x[k];
                    const double d = x[k] - pos0[b][3 * (size_t)i + k];
                    r += d * d;
                }
            }
            double e = 0.0;
            check_rc(ffea_gpu_batch_strain_energy(blobs[b], &e), b, "strain energy");
            strain += e;
            const double rb = std::sqrt(r / nt[b]);
            rmsd2 += rb * rb * nt[b];
            b_strain[b] = e;
            b_rmsd[b] = rb;
            for (int k = 0; k < 3; ++k) b_cog[3 * (size_t)b + k] = c[k] / nt[b];
            for (int k = 0; k < 3; ++k) cog[k] += c[k] / nt[b] * nt[b];
            nodes += nt[b];
        }
        fprintf(traj, "*\n");
        fprintf(traj, "Conformation Changes:\n");
        for (int b = 0; b < p.num_blobs; ++b) fprintf(traj, "Blob %d: Conformation %d -> Conformation %d\n", b, 0, 0);
        fprintf(traj, "*\n");
        fflush(traj);

        const double rmsd = std::sqrt(rmsd2 / nodes);
        for (double &c : cog) c *= 1.0 / nodes;
        fprintf(meas, "%-14.6e", (int)s * p.dt * kTime);
        if (mass_in_system) fprintf(meas, "%-14.6e", kinetic * kEnergy);
        fprintf(meas, "%-14.6e", strain * kEnergy);
        fprintf(meas, "%-14.6e%-14.6e%-14.6e", cog[0] * kLength, cog[1] * kLength, cog[2] * kLength);
        fprintf(meas, "%-14.6e", rmsd * kLength);
        std::vector<double> cell;
        if (p.calc_springs != 0) cell = spring_cells();
        if (p.calc_springs != 0) fprintf(meas, "%-14.6e", spring_total(cell) * kEnergy);
        fprintf(meas, "\n");
        fflush(meas);
        if (!det) return;
        // the detailed line (World.cpp:4504-4537, Blob::write_measurements_to_file
        // Blob.cpp:1204-1213, World::write_detailed_measurements_to_file World.cpp:4694-4721):
        // time, each blob's block, then each pair's block, whose blank lead is written whether
        // or not the header has pair columns
        fprintf(det, "%-14.6e", (int)s * p.dt * kTime);
        for (int b = 0; b < p.num_blobs; ++b) {
            fprintf(det, "     ");
            if (inertial[b]) fprintf(det, "%-14.6e", b_ke[b] * kEnergy);
            fprintf(det, "%-14.6e", b_strain[b] * kEnergy);
            fprintf(det, "%-14.6e%-14.6e%-14.6e%-14.6e", b_cog[3 * (size_t)b] * kLength,
                    b_cog[3 * (size_t)b + 1] * kLength, b_cog[3 * (size_t)b + 2] * kLength, b_rmsd[b] * kLength);
        }
        const int nb = p.num_blobs;
        for (int i = 0; i < nb; ++i)
            for (int j = i; j < nb; ++j) {
                fprintf(det, "       ");
                if (p.calc_springs != 0 && has_springs(i) && has_springs(j)) {
                    // World::get_spring_field_energy(i, j) (World.cpp:4127-4136)
                    const double e = i == j ? cell[(size_t)i * nb + i]
                                            : cell[(size_t)i * nb + j] + cell[(size_t)j * nb + i];
                    fprintf(det, "%-14.6e", e * kEnergy);
                }
            }
        fprintf(det, "\n");
        fflush(det);
    }

    // restart (World.cpp:302-426): the checkpoint's first thermal stream becomes the stream;
    // each blob starts at its own offset in it
    void read_checkpoint() {
        const std::string text = read_whole(files.checkpoint_in);
        std::vector<std::string> lines;
        size_t a = 0;
        while (a < text.size()) {
            size_t e = text.find('\n', a);
            if (e == std::string::npos) e = text.size();
            lines.push_back(text.substr(a, e - a));
            a = e + 1;
        }
        if (lines.empty()) fail(FFEA_GPU_ERR_IO, "Error reading the number of stress seeds: empty checkpoint");
        int nseeds = 0;
        {
            const std::string last = split_spaces(lines[0]).back();
            char *end = nullptr;
            const long v = strtol(last.c_str(), &end, 10);
            if (end == last.c_str()) fail(FFEA_GPU_ERR_IO, "Error reading the number of stress seeds: stoi");
            nseeds = (int)v;
        }
        if (nseeds < 1 || (int)lines.size() < nseeds + 1) fail(FFEA_GPU_ERR_IO, "ERROR reading seeds");
        uint32_t seed[6];
        for (int i = 1; i <= nseeds; ++i) {
            const std::vector<std::string> v = split_spaces(lines[i]);
            if (v.size() != 6) fail(FFEA_GPU_ERR_IO, "ERROR reading seeds");
            for (int j = 0; j < 6; ++j) {
                char *end = nullptr;
                const unsigned long u = strtoul(v[j].c_str(), &end, 10);
                if (end == v[j].c_str()) fail(FFEA_GPU_ERR_IO, "Error reading seeds as integers: stoul");
                if (i == 1) seed[j] = (uint32_t)u;
            }
            if (i > 1) old_seeds.push_back(lines[i]);
        }
        for (int b = 0; b < p.num_blobs; ++b) {
            uint32_t st[6];
            check_rc(ffea_gpu_advance_rng_state(seed, draws_before[b], st), -1, "checkpoint stream state");
            check_rc(ffea_gpu_blob_set_rng_state(blobs[b], st), b, "setting the restart stream state");
        }
    }

    // restart from the trajectory (World.cpp:680-836): searching back from the end for the
    // fourth asterisk finds the start of the last completely written frame; that frame gives
    // the positions and the step to repeat, and the file is cut back to its start
    void restart_trajectory() {
        const std::string t = read_whole(files.trajectory);
        long long pos = (long long)t.size();
        int found = 0;
        bool single = false;
        // 3 asterisks to the top of the last frame, 2 more per deleted frame, 1 for the
        // conformation changes of the frame resumed from (World.cpp:711)
        const int need = 3 + 2 * frames_to_delete + 1;
        while (found != need) {
            if (pos - 2 < 0) {  // the file's start: only one frame; take the header's asterisk
                single = true;
                size_t q = t.find('*', (size_t)pos);
                if (q == std::string::npos) fail(FFEA_GPU_ERR_IO, "no complete frame in " + files.trajectory);
                pos = (long long)q;
            } else {
                pos -= 2;
            }
            const char c = t[(size_t)pos];
            pos += 1;
            if (c == '*') ++found;
            if (single && found != need) fail(FFEA_GPU_ERR_IO, "no complete frame in " + files.trajectory);
        }
        long long cut = pos;
        if (pos < (long long)t.size() && t[(size_t)pos] == '\n') cut = ++pos;
        const char *base = t.c_str();
        const char *cur = base + pos;
        auto expect = [&](const char *lit) {
            const size_t n = strlen(lit);
            if (strncmp(cur, lit, n) != 0) fail(FFEA_GPU_ERR_IO, std::string("malformed trajectory near '") + lit + "'");
            cur += n;
        };
        auto skip_line = [&]() {
            const char *e = strchr(cur, '\n');
            cur = e ? e + 1 : base + t.size();
        };
        if (!single) {
            expect("Conformation Changes:\n");
            for (int b = 0; b < p.num_blobs; ++b) skip_line();
            expect("*\n");
            cut = cur - base;
        }
        long long rstep = -1;
        for (int b = 0; b < p.num_blobs; ++b) {
            int id = -1, conf = -1, used = 0;
            if (sscanf(cur, "Blob %d, Conformation %d, step %lld\n%n", &id, &conf, &rstep, &used) != 3 || used == 0)
                fail(FFEA_GPU_ERR_IO, "Error reading header info for Blob " + std::to_string(b));
            if (id != b)
                fail(FFEA_GPU_ERR_IO, "Mismatch in trajectory file - found blob id = " + std::to_string(id) +
                                          ", was expecting blob id = " + std::to_string(b));
            cur += used;
            skip_line();  // DYNAMIC
            host.resize(3 * (size_t)nt[b]);
            std::vector<double> rvel(inertial[b] ? 3 * (size_t)nt[b] : 0);
            for (int i = 0; i < nt[b]; ++i) {
                double v[10];
                char *end = const_cast<char *>(cur);
                for (int k = 0; k < 10; ++k) {
                    const char *start = end;
                    v[k] = strtod(start, &end);
                    if (end == start)
                        fail(FFEA_GPU_ERR_IO,
                             "(When restarting) Error reading from trajectory file, for node " + std::to_string(i));
                }
                cur = end;
                while (*cur == ' ' || *cur == '\n' || *cur == '\t' || *cur == '\r') ++cur;
                for (int k = 0; k < 3; ++k) host[3 * (size_t)i + k] = v[k] / kLength;
                if (inertial[b])  // Blob::read_nodes_from_file (Blob.cpp:1002-1010)
                    for (int k = 0; k < 3; ++k) rvel[3 * (size_t)i + k] = v[3 + k] / kVelocity;
            }
            check_rc(ffea_gpu_blob_set_positions(blobs[b], host.data()), b, "setting restart positions");
            if (inertial[b]) check_rc(ffea_gpu_blob_set_velocities(blobs[b], rvel.data()), b, "setting restart velocities");
        }
        step_initial = rstep;
        truncate_to(files.trajectory, cut);

        // the measurement files lose their last line, the step that is repeated, and one more
        // per deleted frame (World.cpp:838-866, detailed: 867-912)
        drop_last_lines(files.measurement, frames_to_delete + 1);
        if (detailed) drop_last_lines(files.detailed, frames_to_delete + 1);
    }

    // from the end back over `lines` newlines (starting before the last character), and
    // cut after the last one found
    static void drop_last_lines(const std::string &path, int lines) {
        const std::string m = read_whole(path);
        long long mp = (long long)m.size();
        for (int found = 0; found < lines;) {
            if (mp - 2 < 0) fail(FFEA_GPU_ERR_IO, "Error when trying to find last frame from file " + path);
            mp -= 2;
            const char c = m[(size_t)mp];
            mp += 1;
            if (c == '\n') ++found;
        }
        truncate_to(path, mp);
    }
};

extern "C" {

int ffea_world_create(const char *path, int precision, int device, ffea_world **out) {
    return ffea_world_create_ex(path, precision, device, 0, 0, out);
}

int ffea_world_create_ex(const char *path, int precision, int device, int flags, int frames_to_delete,
                         ffea_world **out) {
    if (!path || !out || (flags & ~FFEA_WORLD_NO_DETAIL) != 0 || frames_to_delete < 0) {
        ffea_internal_set_error("ffea_world_create: invalid argument");
        return FFEA_GPU_ERR_INVALID;
    }
    *out = nullptr;
    ffea_script *s = nullptr;
    int rc = ffea_script_load(path, &s);
    if (rc != FFEA_GPU_OK) return rc;
    ffea_world *w = new ffea_world();
    w->precision = precision;
    w->detailed = (flags & FFEA_WORLD_NO_DETAIL) == 0;
    w->frames_to_delete = frames_to_delete;
    try {
        ffea_script_get_params(s, &w->p);
        ffea_internal_script_files(s, &w->files);
        if (w->files.detailed.empty()) w->detailed = false;  // no name for it (see ScriptFiles)
        const ffea_script_params &p = w->p;
        // SimulationParams::validate (SimulationParams.cpp:805-1009), the file checks
        if (p.restart != 0 && p.restart != 1)
            fail(FFEA_GPU_ERR_INVALID, "Required: Restart flag, 'restart', must be 0 (false) or 1 (true).");
        if (p.restart == 1 && w->files.checkpoint_in.empty())
            fail(FFEA_GPU_ERR_INVALID, "Checkpoint input file required if restart is required.");
        if (w->files.checkpoint_out == w->files.checkpoint_in)
            fail(FFEA_GPU_ERR_INVALID,
                 "it is not allowed to set up checkpoint_in and checkpoint_out with the same file names");
        if (p.check < 1) fail(FFEA_GPU_ERR_INVALID, "'check' must be at least 1");

        long long before = 0;
        for (int b = 0; b < p.num_blobs; ++b) {
            int nn = 0, ne = 0, npin = 0, nctf = 0, nf = 0;
            ffea_script_blob_sizes(s, b, &nn, &ne, &npin, &nctf, &nf);
            ffea_gpu_blob *h = nullptr;
            rc = ffea_gpu_blob_create_from_script(s, b, precision, device, &h);
            if (rc != FFEA_GPU_OK) fail(rc, ffea_gpu_last_error());
            w->blobs.push_back(h);
            w->nt.push_back(nn);
            w->pos0.emplace_back(3 * (size_t)nn);
            w->check_rc(ffea_gpu_blob_get_positions(h, w->pos0.back().data()), b, "reading initial positions");
            w->draws_before.push_back(before);
            before += 7ll * ne;
            int solver = FFEA_GPU_SOLVER_NOMASS_CG;
            ffea_script_blob_solver(s, b, &solver);
            w->inertial.push_back(solver != FFEA_GPU_SOLVER_NOMASS_CG);
            w->mass_in_system |= w->inertial.back() != 0;
            ffea_world::NodeForces F;
            if (w->inertial.back()) {
                F.lin_node.resize(nctf);
                F.lin_force.resize(3 * (size_t)nctf);
                ffea_script_blob_arrays(s, b, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                        F.lin_node.data(), F.lin_force.data(), nullptr);
                int nr = 0;
                ffea_script_blob_rotational_forces(s, b, &nr, nullptr, nullptr, nullptr, nullptr);
                F.rot_node.resize(nr);
                F.rot_force.resize(nr);
                F.rot_axis.resize(6 * (size_t)nr);
                F.rot_type.resize(2 * (size_t)nr);
                ffea_script_blob_rotational_forces(s, b, nullptr, F.rot_node.data(), F.rot_force.data(),
                                                   F.rot_axis.data(), F.rot_type.data());
            }
            w->ctf.push_back(std::move(F));
        }
        if (p.calc_springs == 1) {
            int ns = 0;
            ffea_script_springs(s, &ns, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
            std::vector<double> k(ns), l(ns);
            std::vector<int> blob(2 * (size_t)ns), conf(2 * (size_t)ns), node(2 * (size_t)ns), active(ns);
            ffea_script_springs(s, &ns, k.data(), l.data(), blob.data(), conf.data(), node.data(), active.data());
            for (int i = 0; i < ns; ++i) {
                ffea_world::Spring sp;
                sp.k = k[i];
                sp.l = l[i];
                for (int j = 0; j < 2; ++j) {
                    sp.blob[j] = blob[2 * i + j];
                    sp.node[j] = node[2 * i + j];
                }
                sp.active = active[i] != 0;
                w->springs.push_back(sp);
            }
        }
        ffea_script_free(s);
        s = nullptr;
        // one blob: whole steps between outputs are replayed as graphs (results identical to
        // the eager path). With several blobs every stage runs for all blobs in turn, as the
        // reference does, so a failure leaves every blob at the failing step's positions.
        if (p.num_blobs == 1) w->check_rc(ffea_gpu_blob_use_graph(w->blobs[0], 1, 0), 0, "graph mode");

        move_aside(w->files.checkpoint_out);
        if (p.restart == 0) {
            move_aside(w->files.measurement);
            if (w->detailed) move_aside(w->files.detailed);
            move_aside(w->files.trajectory);
        }
        if (p.restart == 1) w->read_checkpoint();
        if (!(w->ckpt = fopen(w->files.checkpoint_out.c_str(), "wb")))
            fail(FFEA_GPU_ERR_IO, "failed to open " + w->files.checkpoint_out);
        if (p.restart == 0) {
            if (!(w->traj = fopen(w->files.trajectory.c_str(), "wb")))
                fail(FFEA_GPU_ERR_IO, "failed to open " + w->files.trajectory);
            if (!(w->meas = fopen(w->files.measurement.c_str(), "wb")))
                fail(FFEA_GPU_ERR_IO, "failed to open " + w->files.measurement);
            if (w->detailed && !(w->det = fopen(w->files.detailed.c_str(), "wb")))
                fail(FFEA_GPU_ERR_IO, "failed to open " + w->files.detailed);
            w->write_headers(path);
        } else {
            w->restart_trajectory();
            if (!(w->traj = fopen(w->files.trajectory.c_str(), "ab")))
                fail(FFEA_GPU_ERR_IO, "failed to open " + w->files.trajectory);
            if (!(w->meas = fopen(w->files.measurement.c_str(), "ab")))
                fail(FFEA_GPU_ERR_IO, "failed to open " + w->files.measurement);
            if (w->detailed && !(w->det = fopen(w->files.detailed.c_str(), "ab")))
                fail(FFEA_GPU_ERR_IO, "failed to open " + w->files.detailed);
        }
        w->step = w->step_initial;
    } catch (const WorldError &e) {
        if (s) ffea_script_free(s);
        ffea_internal_set_error(e.msg);
        delete w;
        return e.code;
    }
    *out = w;
    return FFEA_GPU_OK;
}

int ffea_world_run(ffea_world *w, long long max_steps) {
    if (!w) return FFEA_GPU_ERR_INVALID;
    const ffea_script_params &p = w->p;
    long long done = 0;
    try {
        // World::run (World.cpp:1747-2060)
        for (; w->step < p.num_steps + 1 && (max_steps < 0 || done < max_steps); ++w->step, ++done) {
            const bool out = w->step % p.check == 0;
            if (out) w->print_checkpoint();
            w->apply_springs();
            if (!out && p.num_blobs == 1) {
                // a step with no output: forces and positions in one call (one graph launch
                // once the iteration count is known); an inversion stops it before the update
                const int rc = ffea_gpu_blob_step(w->blobs[0], 1);
                if (rc != FFEA_GPU_OK) {
                    const std::string m = ffea_gpu_blob_last_error(w->blobs[0]);
                    if (rc == FFEA_GPU_ERR_INVERSION && w->step != w->step_initial) w->print_outputs(w->step);
                    fail(rc, "step " + std::to_string(w->step) + ": " + m);
                }
                continue;
            }
            for (int b = 0; b < p.num_blobs; ++b) {
                const int rc = ffea_gpu_blob_update_internal_forces(w->blobs[b]);
                if (rc != FFEA_GPU_OK) {
                    const std::string m = ffea_gpu_blob_last_error(w->blobs[b]);
                    // World::die_with_dignity (World.cpp:4841-4855): the failing step is dumped
                    if (w->step != w->step_initial) w->print_outputs(w->step);
                    fail(rc, "step " + std::to_string(w->step) + ", blob " + std::to_string(b) + ": " + m);
                }
            }
            if (out) w->print_outputs(w->step);
            for (int b = 0; b < p.num_blobs; ++b)
                w->check_rc(ffea_gpu_blob_update_positions(w->blobs[b]), b, "update_positions");
        }
    } catch (const WorldError &e) {
        w->err = e.msg;
        ffea_internal_set_error(e.msg);
        return e.code;
    }
    return FFEA_GPU_OK;
}

long long ffea_world_current_step(const ffea_world *w) { return w ? w->step : -1; }

int ffea_world_blob(ffea_world *w, int index, ffea_gpu_blob **blob) {
    if (!w || !blob || index < 0 || index >= (int)w->blobs.size()) return FFEA_GPU_ERR_INVALID;
    *blob = w->blobs[index];
    return FFEA_GPU_OK;
}

void ffea_world_destroy(ffea_world *w) { delete w; }

const char *ffea_world_last_error(const ffea_world *w) { return w ? w->err.c_str() : ffea_gpu_last_error(); }

}  // extern "C"
