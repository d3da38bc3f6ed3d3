// Native step driver of the lid-driven cavity: the CManager / CController
// caller of the reference (CManager.cpp:27-209, CController.cpp:168-1510) over
// the C++ solver surface CLbmSolverHIP<T>, on one GPU.
//
//   cavity_driver --size NX NY NZ [--split PX PY PZ] [--u-lid U] [--re RE]
//                 [--loops N] [--precision f32|f64|f16s|bf16s|f16c] [--dump FILE]
//                 [--fused-x]
//   cavity_driver --config FILE.xml [--precision ...] [--loops N] [--dump FILE]
//                 [--vis-format netcdf|netcdf-global|vtk]
//   cavity_driver --config FILE.xml --print-config
//
// * --config reads the reference's XML configuration (CConfiguration.cpp):
//   size, subdomain split, physics, loops, outputs; the CLbmSolver parameters
//   come from its physics for each rank's subdomain (CManager.cpp:28-58);
//   with visualization on, every rank writes the file of its interior at
//   step 0 and after every step whose counter is a multiple of the rate
//   (CController.cpp:1401-1430): netCDF by default, the reference's active
//   writer (CController.cpp:120-135), one file per rank as its serial build
//   or, with --vis-format netcdf-global, one file for the whole cavity as its
//   parallel build; or VTK, the writer it keeps commented out (--vis-format
//   vtk);
// * --print-config prints the parsed fields and the derived parameters of
//   rank 0 and exits without touching the GPU;
// * with validation on and more than one subdomain, the same configuration is
//   run again as one subdomain and the velocities are compared cell by cell
//   (main.cpp:121-303): a cell differs when |u - u_1| / u_1 > epsilon(T) on
//   some component; the differing cells and their count go to
//   <validation dir>/validation_<ranks>.txt (the reference appends the
//   directory and name twice into one path, and indexes the gathered field
//   with swapped strides, which only matters off the cube; both are fixed);
// * subdomains are laid out like MPI ranks, x fastest (CManager.cpp:34-58),
//   with OBSTACLE on global faces and GHOST_LAYER between subdomains (:91-104);
//   the lid goes on the top layer of the top subdomains (:206-208);
// * computeNextStep alternates alpha / beta by step parity (:1278-1287) and
//   exchanges faces after each half-step in the order x, y, z
//   (CController.cpp:369-570, 807-1008), on the device; with --fused-x (a
//   split along x only) the x faces are exchanged inside the step kernels
//   instead, each subdomain reading its neighbours' fused-pack buffers (the
//   same bits, no copy kernel);
// * the benchmark figures are the reference's (CController.cpp:1382-1490):
//   GLUPS over interior cells and GBytes/s = GLUPS * bytes per cell (19 * 2
//   populations + the flag + 4 outputs when stored); with benchmark on in
//   the configuration every rank's figures also go to
//   <benchmark dir>/benchmark_<ranks>_<rank>.txt;
// * --dump writes the interior velocities [3][z][y][x] and densities [z][y][x]
//   of the whole cavity (type T) to FILE.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <array>

#include "CLbmSolverHIP.hpp"
#include "lbm_config.hpp"
#include "lbm_netcdf.hpp"
#include "lbm_vtk.hpp"

namespace {

struct Options {
    int size[3] = {32, 32, 32};
    int split[3] = {1, 1, 1};
    double u_lid = 0.05, re = 100.0;
    int loops = 100;
    std::string precision = "f32";
    std::string dump;
    std::string config;
    std::string vis_format = "netcdf";
    bool print_config = false;
    bool loops_set = false;
    bool fused_x = false;   // x faces exchanged inside the steps (x-only splits)
};

int precision_code(const std::string& p)
{
    if (p == "f32") return LBM_F32;
    if (p == "f64") return LBM_F64;
    if (p == "f16s") return LBM_F16S;
    if (p == "bf16s") return LBM_BF16S;
    if (p == "f16c") return LBM_F16C;
    throw lbm::Error(LBM_ERR_ARG, "unknown precision " + p);
}

// CLbmSolver inputs: from the XML configuration (rank 0's subdomain, as every
// rank derives the same cell length) or from --u-lid / --re with dx = dt = 1.
lbm_physics physics(const Options& o, const lbm::Configuration* cfg)
{
    lbm_physics phys = {};
    if (cfg) {
        int size[3], origin[3];
        double length[3];
        cfg->subdomain(0, size, origin, length);
        phys.length_x = length[0];
        phys.size_x = size[0];
        phys.length_max = std::max(cfg->domain_length[0],
                                   std::max(cfg->domain_length[1], cfg->domain_length[2]));
        for (int a = 0; a < 3; a++) {
            phys.velocity[a] = cfg->velocity[a];
            phys.acceleration[a] = cfg->acceleration[a];
        }
        phys.viscosity = cfg->viscosity;
        phys.timestep = cfg->timestep;
        phys.max_velocity = cfg->max_velocity;
        phys.max_acceleration = cfg->max_acceleration;
        return phys;
    }
    const int nmax = std::max(o.size[0], std::max(o.size[1], o.size[2]));
    phys.length_x = o.size[0];
    phys.size_x = o.size[0];
    phys.length_max = nmax;
    phys.velocity[0] = o.u_lid;
    phys.viscosity = o.u_lid * nmax / o.re;
    phys.timestep = 1.0;
    phys.max_velocity = 0.1;
    phys.max_acceleration = 0.1;
    return phys;
}

template <typename T>
class Controller {
public:
    Controller(const Options& o, const lbm::Configuration* cfg) : o_(o), cfg_(cfg)
    {
        for (int a = 0; a < 3; a++) {
            if (o.size[a] % o.split[a])
                throw lbm::Error(LBM_ERR_ARG, "domain size must be divisible by the split");
            sub_[a] = o.size[a] / o.split[a];
        }
        const int prec = precision_code(o.precision);
        params_ = lbm::computeParameters(physics(o, cfg), prec);
        const bool outputs = cfg ? cfg->store_outputs() : true;
        for (int k = 0; k < o.split[2]; k++)
            for (int j = 0; j < o.split[1]; j++)
                for (int i = 0; i < o.split[0]; i++) {
                    const int idx[3] = {i, j, k};
                    std::vector<lbm::Flag> bc(6, lbm::GHOST_LAYER);
                    for (int a = 0; a < 3; a++) {
                        if (idx[a] == 0)
                            bc[2 * a] = lbm::OBSTACLE;
                        if (idx[a] == o.split[a] - 1)
                            bc[2 * a + 1] = lbm::OBSTACLE;
                    }
                    solvers_.emplace_back(new lbm::CLbmSolverHIP<T>(
                        0, lbm::Vec3i(sub_[0], sub_[1], sub_[2]), bc, params_, outputs, outputs,
                        prec));
                    if (j == o.split[1] - 1)
                        solvers_.back()->setDrivenCavitySzenario();
                }
        if (o.fused_x)
            setup_fused_x();
    }

    ~Controller()
    {
        if (xbufs_.empty())
            return;
        (void)lbm_sync(solvers_[0]->handle(), nullptr);
        for (auto& s : solvers_) {
            (void)lbm_set_step_unpack(s->handle(), nullptr, nullptr);
            (void)lbm_set_step_pack(s->handle(), nullptr, nullptr);
        }
        for (void* p : xbufs_)
            (void)lbm_device_free(0, p);
    }

    Controller(const Controller&) = delete;
    Controller& operator=(const Controller&) = delete;

    const lbm::SolverParameters& params() const { return params_; }

    lbm::CLbmSolverHIP<T>* at(int i, int j, int k)
    {
        return solvers_[(k * o_.split[1] + j) * o_.split[0] + i].get();
    }

    // the faces in the reference order x, then y, then z; the transfers of
    // one axis (both ways between every neighbour pair, listed once) are
    // independent and go to lbm_exchange_faces together
    void exchange(int phase)
    {
        if (!xfers_ready_) {
            for (int axis = 0; axis < 3; axis++)
                for (int k = 0; k < o_.split[2]; k++)
                    for (int j = 0; j < o_.split[1]; j++)
                        for (int i = 0; i < o_.split[0]; i++) {
                            int nb[3] = {i, j, k};
                            nb[axis]++;
                            if (nb[axis] >= o_.split[axis])
                                continue;
                            auto* a = at(i, j, k);
                            auto* b = at(nb[0], nb[1], nb[2]);
                            xfers_[axis].push_back({a->handle(), 2 * axis + 1, b->handle()});
                            xfers_[axis].push_back({b->handle(), 2 * axis, a->handle()});
                        }
            xfers_ready_ = true;
        }
        for (const auto& t : xfers_)
            lbm::check(lbm_exchange_faces(t.data(), (int)t.size(), phase, nullptr));
    }

    // x faces exchanged inside the step kernels: every subdomain packs its
    // communicating x faces into buffers of its own (lbm_set_step_pack) and
    // its steps read the neighbours' packs as received faces
    // (lbm_set_step_unpack); the packs after alpha and after beta carry
    // disjoint planes, so one buffer per face serves both. Exact for splits
    // along x only (with y / z splits the reference's y / z exchange forwards
    // received x values to edge cells before a fused unpack would apply them).
    void setup_fused_x()
    {
        if (o_.split[0] < 2 || o_.split[1] != 1 || o_.split[2] != 1)
            throw lbm::Error(LBM_ERR_ARG, "--fused-x needs a split along x only");
        const int n = o_.split[0];
        std::vector<void*> left(n, nullptr), right(n, nullptr);
        for (int i = 0; i < n; i++) {
            lbm_solver* h = at(i, 0, 0)->handle();
            for (int d = 0; d < 2; d++) {
                if ((d == 0 && i == 0) || (d == 1 && i == n - 1))
                    continue;
                void* p = nullptr;
                lbm::check(lbm_device_alloc(0, (size_t)lbm_face_elements(h, d) * sizeof(T), &p));
                xbufs_.push_back(p);
                (d ? right : left)[i] = p;
            }
            at(i, 0, 0)->setStepPack(static_cast<T*>(left[i]), static_cast<T*>(right[i]));
        }
        for (int i = 0; i < n; i++)
            xrecv_.push_back({i > 0 ? right[i - 1] : nullptr, i < n - 1 ? left[i + 1] : nullptr});
    }

    void computeNextStep()
    {
        const int phase = counter_ & 1;
        if (phase && !xrecv_.empty() && !xarmed_)
            throw lbm::Error(LBM_ERR_ARG, "--fused-x: the first step must be an alpha step");
        for (auto& s : solvers_) {
            if (phase)
                s->simulationStepBeta();
            else
                s->simulationStepAlpha();
        }
        if (!xrecv_.empty()) {
            // from the first beta on: before it nothing was exchanged
            if (!xarmed_) {
                for (size_t i = 0; i < solvers_.size(); i++)
                    solvers_[i]->setStepUnpack(static_cast<const T*>(xrecv_[i][0]),
                                               static_cast<const T*>(xrecv_[i][1]));
                xarmed_ = true;
            }
        } else if (solvers_.size() > 1) {
            exchange(phase);
        }
        counter_++;
    }

    void sync() { lbm::check(lbm_sync(solvers_[0]->handle(), nullptr)); }

    long counter() const { return counter_; }

    // CLbmVisualization::render of every rank (CController.cpp:1401-1430)
    void render(long iteration)
    {
        if (!cfg_ || !cfg_->do_visualization || iteration % cfg_->visualization_rate)
            return;
        if (o_.vis_format == "netcdf-global") {
            std::vector<T> u, rho;
            std::vector<int32_t> flags;
            gather(u, rho, &flags);
            int size[3], origin[3];
            double length[3];
            cfg_->subdomain(0, size, origin, length);
            const T len[3] = {(T)length[0], (T)length[1], (T)length[2]};
            lbm::write_netcdf_global<T>(cfg_->visualization_dir, (int)iteration, o_.size, len,
                                        size, flags.data(), rho.data(), u.data());
            return;
        }
        const long sc = (long)sub_[0] * sub_[1] * sub_[2];
        std::vector<T> rho(sc), u(3 * sc);
        std::vector<lbm::Flag> flags(sc);
        std::vector<int32_t> flag_values(sc);
        for (int rank = 0; rank < (int)solvers_.size(); rank++) {
            int size[3], origin[3];
            double length[3];
            cfg_->subdomain(rank, size, origin, length);
            const T len[3] = {(T)length[0], (T)length[1], (T)length[2]};
            auto* s = solvers_[rank].get();
            s->getFlags(flags.data());
            s->getDensities(rho.data());
            s->getVelocities(u.data());
            for (long c = 0; c < sc; c++)
                flag_values[c] = flags[c];
            if (o_.vis_format == "vtk")
                lbm::write_vtk<T>(cfg_->visualization_dir, rank, (int)iteration, size, origin,
                                  len, flag_values.data(), rho.data(), u.data());
            else
                lbm::write_netcdf<T>(cfg_->visualization_dir, rank, (int)iteration, size, len,
                                     flag_values.data(), rho.data(), u.data());
        }
    }

    // interior velocities [3][z][y][x] of the whole cavity
    std::vector<T> velocities()
    {
        std::vector<T> u, rho;
        gather(u, rho);
        return u;
    }

    void gather(std::vector<T>& u, std::vector<T>& rho, std::vector<int32_t>* flags = nullptr)
    {
        const long nx = o_.size[0], ny = o_.size[1], nz = o_.size[2], cells = nx * ny * nz;
        u.assign(3 * cells, T(0));
        rho.assign(cells, T(0));
        if (flags)
            flags->assign(cells, 0);
        const long sc = (long)sub_[0] * sub_[1] * sub_[2];
        std::vector<T> pu(3 * sc), pr(sc);
        std::vector<lbm::Flag> pf(flags ? sc : 0);
        for (int k = 0; k < o_.split[2]; k++)
            for (int j = 0; j < o_.split[1]; j++)
                for (int i = 0; i < o_.split[0]; i++) {
                    auto* s = at(i, j, k);
                    s->getVelocities(pu.data());
                    s->getDensities(pr.data());
                    if (flags)
                        s->getFlags(pf.data());
                    for (int z = 0; z < sub_[2]; z++)
                        for (int y = 0; y < sub_[1]; y++)
                            for (int x = 0; x < sub_[0]; x++) {
                                const long lid = ((long)z * sub_[1] + y) * sub_[0] + x;
                                const long gid = ((long)(k * sub_[2] + z) * ny + (j * sub_[1] + y)) * nx +
                                                 (i * sub_[0] + x);
                                for (int c = 0; c < 3; c++)
                                    u[c * cells + gid] = pu[c * sc + lid];
                                rho[gid] = pr[lid];
                                if (flags)
                                    (*flags)[gid] = pf[lid];
                            }
                }
    }

    void dump(const std::string& path)
    {
        std::vector<T> u, rho;
        gather(u, rho);
        std::ofstream f(path, std::ios::binary);
        f.write(reinterpret_cast<const char*>(u.data()), u.size() * sizeof(T));
        f.write(reinterpret_cast<const char*>(rho.data()), rho.size() * sizeof(T));
        if (!f)
            throw lbm::Error(LBM_ERR_ARG, "cannot write " + path);
    }

private:
    Options o_;
    const lbm::Configuration* cfg_;
    int sub_[3];
    lbm::SolverParameters params_;
    std::vector<std::unique_ptr<lbm::CLbmSolverHIP<T>>> solvers_;
    std::vector<void*> xbufs_;                  // fused x exchange: pack buffers
    std::vector<std::array<void*, 2>> xrecv_;   // per solver: received LEFT, RIGHT
    bool xarmed_ = false;
    std::array<std::vector<lbm_face_transfer>, 3> xfers_;   // per axis
    bool xfers_ready_ = false;
    long counter_ = 0;
};

// main.cpp:121-303: the configuration again on one subdomain (the
// reference's serialize(); its visualization is not repeated here), then a
// cell-by-cell comparison of the velocities.
template <typename T>
int validate(const Options& o, const lbm::Configuration& cfg, const std::vector<T>& u, int ranks)
{
    lbm::Configuration one = cfg;
    for (int a = 0; a < 3; a++)
        one.subdomains[a] = 1;
    one.do_visualization = 0;
    Options o1 = o;
    for (int a = 0; a < 3; a++)
        o1.split[a] = 1;
    Controller<T> mono(o1, &one);
    for (int i = 0; i < o.loops; i++)
        mono.computeNextStep();
    mono.sync();
    const std::vector<T> v = mono.velocities();
    const long nx = o.size[0], ny = o.size[1], nz = o.size[2], cells = nx * ny * nz;
    std::stringstream name;
    name << cfg.validation_dir << "/validation_" << ranks << ".txt";
    std::ofstream f(name.str().c_str(), std::ios::out);
    if (!f)
        throw lbm::Error(LBM_ERR_ARG, "cannot open " + name.str());
    const T eps = std::numeric_limits<T>::epsilon();
    long unequal = 0;
    for (long i = 0; i < nx; i++)
        for (long j = 0; j < ny; j++)
            for (long k = 0; k < nz; k++) {
                const long g = (k * ny + j) * nx + i;
                bool diff = false;
                for (int c = 0; c < 3; c++)
                    diff = diff || std::fabs(u[c * cells + g] - v[c * cells + g]) / v[c * cells + g] > eps;
                if (!diff)
                    continue;
                f << "[" << i << "," << j << "," << k << "]: [" << u[g] << ", " << u[cells + g] << ", "
                  << u[2 * cells + g] << "] vs. [" << v[g] << ", " << v[cells + g] << ", "
                  << v[2 * cells + g] << "]" << std::endl;
                unequal++;
            }
    f << "number of inequalities: " << unequal << std::endl;
    std::printf("validation:      %ld inequalities (%s)\n", unequal, name.str().c_str());
    return 0;
}

template <typename T>
int run(const Options& o, const lbm::Configuration* cfg)
{
    Controller<T> c(o, cfg);
    std::printf("tau:             %.17g\ninv tau:         %.17g\n", c.params().tau, c.params().tauInv);
    c.render(0);
    c.sync();
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < o.loops; i++) {
        c.computeNextStep();
        c.render(c.counter());
    }
    c.sync();
    const double elapsed =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // CController.cpp:1380-1399, 1451-1466: bytes per cell = populations read
    // and written, the flag, and the stored rho / u when there are outputs;
    // the figures are computed in T and printed with stream defaults
    const int pop_bytes = o.precision == "f64" ? 8 : (o.precision == "f32" ? 4 : 2);
    const bool outputs = cfg ? cfg->store_outputs() : true;
    const int used = 19 * 2 * pop_bytes + 4 + (outputs ? 4 * (int)sizeof(T) : 0);
    const T el = (T)elapsed;
    const T its = (T)o.loops / el;
    const T glups = its * (T)o.size[0] * (T)o.size[1] * (T)o.size[2] * (T)0.000000001;
    const T bandwidth = glups * (T)used;
    std::stringstream report;
    report << "loops:           " << o.loops << std::endl
           << "time:            " << el << "s" << std::endl
           << "iterations:      " << its << "s^-1" << std::endl
           << "lattice updates: " << glups << "GLUPS" << std::endl
           << "bandwidth:       " << bandwidth << "GBytes/s" << std::endl;
    std::fputs(report.str().c_str(), stdout);
    const int ranks = o.split[0] * o.split[1] * o.split[2];
    if (cfg && cfg->do_benchmark)
        for (int rank = 0; rank < ranks; rank++) {
            std::stringstream name;
            name << cfg->benchmark_dir << "/benchmark_" << ranks << "_" << rank << ".txt";
            std::ofstream f(name.str().c_str(), std::ios::out);
            if (!f)
                throw lbm::Error(LBM_ERR_ARG, "cannot open " + name.str());
            f << report.str();
        }
    if (!o.dump.empty())
        c.dump(o.dump);
    if (cfg && cfg->do_validation && ranks > 1)
        return validate<T>(o, *cfg, c.velocities(), ranks);
    return 0;
}

// Parsed fields (in the order velocity, acceleration, viscosity, max velocity /
// acceleration, size, length, subdomains, CPU ratio, loops, timestep, four
// switches, visualization rate, CPU blocks, GPU grids) and rank 0's
// CLbmSolver parameters, one "key: values" line each, %.17g.
void print_config(const lbm::Configuration& c, const lbm::SolverParameters& p)
{
    std::vector<double> v;
    for (int a = 0; a < 3; a++) v.push_back(c.velocity[a]);
    for (int a = 0; a < 3; a++) v.push_back(c.acceleration[a]);
    v.push_back(c.viscosity);
    v.push_back(c.max_velocity);
    v.push_back(c.max_acceleration);
    for (int a = 0; a < 3; a++) v.push_back(c.domain_size[a]);
    for (int a = 0; a < 3; a++) v.push_back(c.domain_length[a]);
    for (int a = 0; a < 3; a++) v.push_back(c.subdomains[a]);
    v.push_back(c.cpu_subdomain_ratio);
    v.push_back(c.loops);
    v.push_back(c.timestep);
    v.push_back(c.do_benchmark);
    v.push_back(c.do_logging);
    v.push_back(c.do_validation);
    v.push_back(c.do_visualization);
    v.push_back(c.visualization_rate);
    for (auto& b : c.cpu_blocks)
        for (int a = 0; a < 3; a++) v.push_back(b[a]);
    for (auto& g : c.gpu_grids)
        for (int a = 0; a < 3; a++) v.push_back(g[a]);
    std::printf("fields:");
    for (double x : v)
        std::printf(" %.17g", x);
    std::printf("\nvisualization-dir: %s\n", c.visualization_dir.c_str());
    std::printf("params: %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                p.velocityDimLess[0], p.velocityDimLess[1], p.velocityDimLess[2],
                p.accelerationDimLess[0], p.accelerationDimLess[1], p.accelerationDimLess[2],
                p.viscosityDimLess, p.tau, p.tauInv, p.timestep, p.viscosity);
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto need = [&](int n) {
            if (i + n >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
        };
        if (a == "--size") {
            need(3);
            for (int k = 0; k < 3; k++) o.size[k] = std::atoi(argv[++i]);
        } else if (a == "--split") {
            need(3);
            for (int k = 0; k < 3; k++) o.split[k] = std::atoi(argv[++i]);
        } else if (a == "--u-lid") {
            need(1);
            o.u_lid = std::atof(argv[++i]);
        } else if (a == "--re") {
            need(1);
            o.re = std::atof(argv[++i]);
        } else if (a == "--loops") {
            need(1);
            o.loops = std::atoi(argv[++i]);
            o.loops_set = true;
        } else if (a == "--config") {
            need(1);
            o.config = argv[++i];
        } else if (a == "--print-config") {
            o.print_config = true;
        } else if (a == "--fused-x") {
            o.fused_x = true;
        } else if (a == "--precision") {
            need(1);
            o.precision = argv[++i];
        } else if (a == "--dump") {
            need(1);
            o.dump = argv[++i];
        } else if (a == "--vis-format") {
            need(1);
            o.vis_format = argv[++i];
            if (o.vis_format != "netcdf" && o.vis_format != "netcdf-global" &&
                o.vis_format != "vtk") {
                std::fprintf(stderr, "--vis-format takes netcdf, netcdf-global or vtk\n");
                return 2;
            }
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    try {
        std::unique_ptr<lbm::Configuration> cfg;
        if (!o.config.empty()) {
            cfg.reset(new lbm::Configuration(lbm::load_configuration(o.config, o.precision == "f64")));
            for (int a = 0; a < 3; a++) {
                o.size[a] = cfg->domain_size[a];
                o.split[a] = cfg->subdomains[a];
            }
            if (!o.loops_set)
                o.loops = cfg->loops;
        }
        if (o.print_config) {
            if (!cfg)
                throw lbm::Error(LBM_ERR_ARG, "--print-config needs --config");
            print_config(*cfg, lbm::computeParameters(physics(o, cfg.get()),
                                                      precision_code(o.precision)));
            return 0;
        }
        if (o.loops < 0)
            throw lbm::Error(LBM_ERR_ARG, "loops < 0 (an endless run) is not offered");
        if (o.precision == "f64")
            return run<double>(o, cfg.get());
        return run<float>(o, cfg.get());
    } catch (const lbm::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    } catch (const std::runtime_error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
