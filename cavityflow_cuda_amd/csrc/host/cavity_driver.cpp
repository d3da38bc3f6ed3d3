// Native step driver of the lid-driven cavity: the CManager / CController
// caller of the reference (CManager.cpp:27-209, CController.cpp:168-1510) over
// the C++ solver surface CLbmSolverHIP<T>, on one GPU.
//
//   cavity_driver --size NX NY NZ [--split PX PY PZ] [--u-lid U] [--re RE]
//                 [--loops N] [--precision f32|f64|f16s|bf16s] [--dump FILE]
//
// * subdomains are laid out like MPI ranks, x fastest (CManager.cpp:34-58),
//   with OBSTACLE on global faces and GHOST_LAYER between subdomains (:91-104);
//   the lid goes on the top layer of the top subdomains (:206-208);
// * computeNextStep alternates alpha / beta by step parity (:1278-1287) and
//   exchanges faces after each half-step in the order x, y, z
//   (CController.cpp:369-570, 807-1008), on the device;
// * the benchmark figures are the reference's (CController.cpp:1382-1490):
//   GLUPS over interior cells and GBytes/s = GLUPS * (19 * 2 * sizeof(T) + 4);
// * --dump writes the interior velocities [3][z][y][x] and densities [z][y][x]
//   of the whole cavity (type T) to FILE.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "CLbmSolverHIP.hpp"

namespace {

struct Options {
    int size[3] = {32, 32, 32};
    int split[3] = {1, 1, 1};
    double u_lid = 0.05, re = 100.0;
    int loops = 100;
    std::string precision = "f32";
    std::string dump;
};

int precision_code(const std::string& p)
{
    if (p == "f32") return LBM_F32;
    if (p == "f64") return LBM_F64;
    if (p == "f16s") return LBM_F16S;
    if (p == "bf16s") return LBM_BF16S;
    throw lbm::Error(LBM_ERR_ARG, "unknown precision " + p);
}

template <typename T>
class Controller {
public:
    Controller(const Options& o) : o_(o)
    {
        for (int a = 0; a < 3; a++) {
            if (o.size[a] % o.split[a])
                throw lbm::Error(LBM_ERR_ARG, "domain size must be divisible by the split");
            sub_[a] = o.size[a] / o.split[a];
        }
        const int nmax = std::max(o.size[0], std::max(o.size[1], o.size[2]));
        lbm_physics phys = {};
        phys.length_x = o.size[0];
        phys.size_x = o.size[0];
        phys.length_max = nmax;
        phys.velocity[0] = o.u_lid;
        phys.viscosity = o.u_lid * nmax / o.re;
        phys.timestep = 1.0;
        phys.max_velocity = 0.1;
        phys.max_acceleration = 0.1;
        const int prec = precision_code(o.precision);
        params_ = lbm::computeParameters(phys, prec);
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
                        0, lbm::Vec3i(sub_[0], sub_[1], sub_[2]), bc, params_, true, true, prec));
                    if (j == o.split[1] - 1)
                        solvers_.back()->setDrivenCavitySzenario();
                }
    }

    const lbm::SolverParameters& params() const { return params_; }

    lbm::CLbmSolverHIP<T>* at(int i, int j, int k)
    {
        return solvers_[(k * o_.split[1] + j) * o_.split[0] + i].get();
    }

    void exchange(int phase)
    {
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
                        lbm::check(lbm_exchange_face(a->handle(), 2 * axis + 1, phase, b->handle(),
                                                     nullptr));
                        lbm::check(lbm_exchange_face(b->handle(), 2 * axis, phase, a->handle(),
                                                     nullptr));
                    }
    }

    void computeNextStep()
    {
        const int phase = counter_ & 1;
        for (auto& s : solvers_) {
            if (phase)
                s->simulationStepBeta();
            else
                s->simulationStepAlpha();
        }
        if (solvers_.size() > 1)
            exchange(phase);
        counter_++;
    }

    void sync() { lbm::check(lbm_sync(solvers_[0]->handle(), nullptr)); }

    void dump(const std::string& path)
    {
        const long nx = o_.size[0], ny = o_.size[1], nz = o_.size[2], cells = nx * ny * nz;
        std::vector<T> u(3 * cells), rho(cells);
        const long sc = (long)sub_[0] * sub_[1] * sub_[2];
        std::vector<T> pu(3 * sc), pr(sc);
        for (int k = 0; k < o_.split[2]; k++)
            for (int j = 0; j < o_.split[1]; j++)
                for (int i = 0; i < o_.split[0]; i++) {
                    auto* s = at(i, j, k);
                    s->getVelocities(pu.data());
                    s->getDensities(pr.data());
                    for (int z = 0; z < sub_[2]; z++)
                        for (int y = 0; y < sub_[1]; y++)
                            for (int x = 0; x < sub_[0]; x++) {
                                const long lid = ((long)z * sub_[1] + y) * sub_[0] + x;
                                const long gid = ((long)(k * sub_[2] + z) * ny + (j * sub_[1] + y)) * nx +
                                                 (i * sub_[0] + x);
                                for (int c = 0; c < 3; c++)
                                    u[c * cells + gid] = pu[c * sc + lid];
                                rho[gid] = pr[lid];
                            }
                }
        std::ofstream f(path, std::ios::binary);
        f.write(reinterpret_cast<const char*>(u.data()), u.size() * sizeof(T));
        f.write(reinterpret_cast<const char*>(rho.data()), rho.size() * sizeof(T));
        if (!f)
            throw lbm::Error(LBM_ERR_ARG, "cannot write " + path);
    }

private:
    Options o_;
    int sub_[3];
    lbm::SolverParameters params_;
    std::vector<std::unique_ptr<lbm::CLbmSolverHIP<T>>> solvers_;
    long counter_ = 0;
};

template <typename T>
int run(const Options& o)
{
    Controller<T> c(o);
    std::printf("tau:             %.17g\ninv tau:         %.17g\n", c.params().tau, c.params().tauInv);
    c.sync();
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < o.loops; i++)
        c.computeNextStep();
    c.sync();
    const double elapsed =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double its = o.loops / elapsed;
    const double glups = its * o.size[0] * (double)o.size[1] * o.size[2] * 1e-9;
    const int pop_bytes = o.precision == "f64" ? 8 : (o.precision == "f32" ? 4 : 2);
    std::printf("loops:           %d\ntime:            %gs\niterations:      %gs^-1\n"
                "lattice updates: %gGLUPS\nbandwidth:       %gGBytes/s\n",
                o.loops, elapsed, its, glups, glups * (19 * 2 * pop_bytes + 4));
    if (!o.dump.empty())
        c.dump(o.dump);
    return 0;
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
        } else if (a == "--precision") {
            need(1);
            o.precision = argv[++i];
        } else if (a == "--dump") {
            need(1);
            o.dump = argv[++i];
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    try {
        if (o.precision == "f64")
            return run<double>(o);
        return run<float>(o);
    } catch (const lbm::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
