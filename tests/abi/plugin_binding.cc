// plugin_binding.cc -- the v2RDMSolver side of the binding as a maintainer
// would add it to the reference plugin (v2rdm_solver.cc), compiled as plain
// C++ against include/v2g.h and linked with libv2g.so.
//
// PluginSide keeps the reference's member layout (host FP64 vectors x, z, y,
// c, b, one per SharedVector) and its seams' signatures:
//   bpsdp_Au(A, u), bpsdp_ATu(A, u)   v2rdm_solver.cc:3020-3065, 3114-3158
//   cg_Ax(N, A, u)                    v2rdm_solver.cc:3207-3213
//   Update_xz()                       v2rdm_solver.cc:3216-3289
//   compute_energy()                  v2rdm_solver.cc:1778-1912 (whole loop)
// each forwarding to the engine on device copies.
//
// Usage: plugin_binding nirrep amopi[nirrep] nalpha nbeta ints.bin out.bin
//   ints.bin: amo*amo h1 then amo^4 (pq|rs) doubles (Pitzer order, row-major)
//   out.bin:  x after the solve (dimx), then A u and A^T w for the fixed test
//             vectors u_i = sin(i + 1) (dimx) and w_i = cos(i + 1) (ncon)
// Prints one JSON line with the sizes, the energy and the iteration count.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/v2g.h"

namespace {

void chk(int rc, const char *what)
{
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + v2g_last_error());
}

void hchk(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct DeviceVec {
    double *p = nullptr;
    long long n = 0;
    explicit DeviceVec(long long count) : n(count) { hchk(hipMalloc(&p, 8 * n), "hipMalloc"); }
    ~DeviceVec() { (void)hipFree(p); }
    void put(const std::vector<double> &v) { hchk(hipMemcpy(p, v.data(), 8 * n, hipMemcpyHostToDevice), "H2D"); }
    void get(std::vector<double> &v) const { hchk(hipMemcpy(v.data(), p, 8 * n, hipMemcpyDeviceToHost), "D2H"); }
};

class PluginSide {
public:
    PluginSide(const v2g_layout_desc &d)
    {
        chk(v2g_create(&d, &gpu_), "v2g_create");
        int nblk = 0;
        chk(v2g_dims(gpu_, &dimx_, &ncon_, &nblk), "v2g_dims");
        x_.resize(dimx_); z_.resize(dimx_); c_.resize(dimx_);
        y_.resize(ncon_); b_.resize(ncon_);
        d_x_ = new DeviceVec(dimx_); d_z_ = new DeviceVec(dimx_); d_c_ = new DeviceVec(dimx_);
        d_y_ = new DeviceVec(ncon_); d_b_ = new DeviceVec(ncon_);
    }
    ~PluginSide()
    {
        delete d_x_; delete d_z_; delete d_c_; delete d_y_; delete d_b_;
        v2g_destroy(gpu_);
    }
    long long dimx() const { return dimx_; }
    long long ncon() const { return ncon_; }
    const std::vector<double> &x() const { return x_; }

    // BuildConstraints / GetIntegrals (RepackIntegrals) / Guess
    void setup(const std::vector<double> &h1, const std::vector<double> &eri)
    {
        chk(v2g_build_b(gpu_, b_.data()), "v2g_build_b");
        chk(v2g_build_c(gpu_, h1.data(), eri.data(), c_.data()), "v2g_build_c");
        chk(v2g_guess_random(gpu_, x_.data(), z_.data(), y_.data()), "v2g_guess_random");
    }

    // void bpsdp_Au(SharedVector A, SharedVector u): A = A(u), callee zeroes A
    void bpsdp_Au(std::vector<double> &A, const std::vector<double> &u)
    {
        DeviceVec du(dimx_), dA(ncon_);
        du.put(u);
        chk(v2g_Au(gpu_, du.p, dA.p), "v2g_Au");
        dA.get(A);
    }
    void bpsdp_ATu(std::vector<double> &A, const std::vector<double> &u)
    {
        DeviceVec du(ncon_), dA(dimx_);
        du.put(u);
        chk(v2g_ATu(gpu_, du.p, dA.p), "v2g_ATu");
        dA.get(A);
    }
    void cg_Ax(long long N, std::vector<double> &A, const std::vector<double> &u)
    {
        if (N != ncon_) throw std::runtime_error("dimension does not match dimension from initialization");
        DeviceVec du(ncon_), dA(ncon_);
        du.put(u);
        chk(v2g_cg_Ax(gpu_, du.p, dA.p), "v2g_cg_Ax");
        dA.get(A);
    }
    void Update_xz(double mu)
    {
        d_x_->put(x_); d_y_->put(y_); d_c_->put(c_);
        chk(v2g_update_xz(gpu_, mu, d_y_->p, d_c_->p, d_x_->p, d_z_->p), "v2g_update_xz");
        d_x_->get(x_); d_z_->get(z_);
    }

    // the do { } while loop of compute_energy, device-resident
    v2g_stats compute_energy()
    {
        d_x_->put(x_); d_z_->put(z_); d_y_->put(y_); d_c_->put(c_); d_b_->put(b_);
        v2g_params p{};
        p.tau = 1.0; p.r_convergence = 1e-6; p.e_convergence = 1e-6; p.cg_convergence = 1e-9;
        p.maxiter = 100000; p.cg_maxiter = 10000; p.mu_update_frequency = 500;
        p.precond = 0; p.cg_extrapolate = 1; p.eig_warm_sweeps = 1; p.eig_inner_sweeps = 1;
        v2g_stats st{};
        chk(v2g_bpsdp(gpu_, &p, d_c_->p, d_b_->p, d_x_->p, d_z_->p, d_y_->p, &st), "v2g_bpsdp");
        if (!st.converged) throw std::runtime_error("v2RDM did not converge.");
        d_x_->get(x_); d_z_->get(z_); d_y_->get(y_);
        return st;
    }

private:
    v2g_handle gpu_ = nullptr;
    long long dimx_ = 0, ncon_ = 0;
    std::vector<double> x_, z_, y_, c_, b_;
    DeviceVec *d_x_, *d_z_, *d_c_, *d_y_, *d_b_;
};

}  // namespace

int main(int argc, char **argv)
{
    try {
        if (argc < 2) throw std::runtime_error("usage: plugin_binding nirrep amopi... na nb ints out");
        v2g_layout_desc d{};
        d.nirrep = std::atoi(argv[1]);
        if (argc != 2 + d.nirrep + 4) throw std::runtime_error("wrong argument count");
        int amo = 0;
        for (int h = 0; h < d.nirrep; h++) amo += (d.amopi[h] = std::atoi(argv[2 + h]));
        d.nalpha = std::atoi(argv[2 + d.nirrep]);
        d.nbeta = std::atoi(argv[3 + d.nirrep]);
        d.multiplicity = d.nalpha - d.nbeta + 1;
        d.constrain_q2 = d.constrain_g2 = 1;
        d.constrain_spin = d.constrain_sz = 1;
        std::vector<double> h1((size_t)amo * amo), eri((size_t)amo * amo * amo * amo);
        FILE *f = std::fopen(argv[4 + d.nirrep], "rb");
        if (!f) throw std::runtime_error("cannot open the integral file");
        const bool ok = std::fread(h1.data(), 8, h1.size(), f) == h1.size() &&
                        std::fread(eri.data(), 8, eri.size(), f) == eri.size();
        std::fclose(f);
        if (!ok) throw std::runtime_error("short integral file");

        PluginSide s(d);
        s.setup(h1, eri);
        const v2g_stats st = s.compute_energy();
        std::vector<double> u(s.dimx()), w(s.ncon()), Au(s.ncon()), ATw(s.dimx()), AATw(s.ncon());
        for (long long i = 0; i < s.dimx(); i++) u[i] = std::sin((double)(i + 1));
        for (long long i = 0; i < s.ncon(); i++) w[i] = std::cos((double)(i + 1));
        s.bpsdp_Au(Au, u);
        s.bpsdp_ATu(ATw, w);
        s.cg_Ax(s.ncon(), AATw, w);
        double chk_aat = 0.0;   // <w, A A^T w> = ||A^T w||^2
        double nrm = 0.0;
        for (long long i = 0; i < s.ncon(); i++) chk_aat += w[i] * AATw[i];
        for (long long i = 0; i < s.dimx(); i++) nrm += ATw[i] * ATw[i];

        FILE *o = std::fopen(argv[5 + d.nirrep], "wb");
        if (!o) throw std::runtime_error("cannot open the output file");
        std::fwrite(s.x().data(), 8, s.x().size(), o);
        std::fwrite(Au.data(), 8, Au.size(), o);
        std::fwrite(ATw.data(), 8, ATw.size(), o);
        std::fclose(o);
        std::printf("{\"dimx\": %lld, \"ncon\": %lld, \"converged\": %d, \"oiter\": %d, "
                    "\"energy\": %.12f, \"cg_Ax_consistency\": %.3e}\n",
                    s.dimx(), s.ncon(), st.converged, st.oiter, st.energy_primal,
                    std::fabs(chk_aat - nrm) / nrm);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "plugin_binding: %s\n", e.what());
        return 1;
    }
}
