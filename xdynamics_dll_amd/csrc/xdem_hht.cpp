// xdem_hht.cpp -- HHT-alpha integrator of a driven wheel body (host side of the DEM <-> MBD
// coupling; include/xdem_mbd.h).
//
// Restates the reference's implicit multibody step (xIntegratorHHT.cpp:50-202 with the
// assembly of xMultiBodySimulation.cpp:291-440) for one body in Euler parameters,
// q = (x, y, z, e0, e1, e2, e3), and the constraint set of a single-wheel rig:
//   phi0 = y - y0                                  (the wheel stays in its plane)
//   phi1 = A(e)[0][1], phi2 = A(e)[2][1]            (its y axis stays the global y axis)
//   phi3 = e2 cos(theta/2) - e0 sin(theta/2)        (rotation about y: theta = theta0 + omega t)
//   phi4 = e . e - 1                                (Euler-parameter norm)
// Mass matrix diag(m) + 4 G^T J G; generalised force (m g + load + F_contact,
// 2 G^T A^T M_contact + 8 G(e_dot)^T J G(e_dot) e); HHT prediction, Newton correction with the
// constraint equations scaled by -1 / (beta dt^2) and the mass Jacobian as the reference
// assembles them. Differences, by choice: alpha, beta and gamma all follow the configured
// alpha (the reference's Initialize resets alpha to -0.3 after any earlier setting, and a later
// setAlphaValue changes alpha alone, leaving beta, gamma and 1 / (1 + alpha) at -0.3's values);
// no parabolic predictor of the Newton start (it changes the iterations, not the converged
// step); a dense LU with partial pivoting in place of LAPACKE's dgesv.
#include <cmath>
#include <cstring>
#include <new>

#include "xdem_mbd.h"
#include "xdem_types.h"

namespace {

constexpr int kQ = 7;          // coordinates
constexpr int kC = 5;          // constraints
constexpr int kT = kQ + kC;    // unknowns of the Newton system: accelerations and multipliers

struct Mat34 { double a[3][4]; };

// G(e) (xAlgebraMath.cpp:663-670)
Mat34 gmat(const double* e)
{
    Mat34 g = {{{-e[1], e[0], e[3], -e[2]}, {-e[2], -e[3], e[0], e[1]}, {-e[3], e[2], -e[1], e[0]}}};
    return g;
}

// Rotation matrix A(e) (xPointMass.cpp:296-301)
void amat(const double* e, double A[3][3])
{
    A[0][0] = 2 * (e[0] * e[0] + e[1] * e[1] - 0.5); A[0][1] = 2 * (e[1] * e[2] - e[0] * e[3]); A[0][2] = 2 * (e[1] * e[3] + e[0] * e[2]);
    A[1][0] = 2 * (e[1] * e[2] + e[0] * e[3]); A[1][1] = 2 * (e[0] * e[0] + e[2] * e[2] - 0.5); A[1][2] = 2 * (e[2] * e[3] - e[0] * e[1]);
    A[2][0] = 2 * (e[1] * e[3] - e[0] * e[2]); A[2][1] = 2 * (e[2] * e[3] + e[0] * e[1]); A[2][2] = 2 * (e[0] * e[0] + e[3] * e[3] - 0.5);
}

// Solves M x = b in place (n <= kT), LU with partial pivoting; returns false when singular.
bool solve(double M[kT][kT], double* b, int n)
{
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int r = k + 1; r < n; r++)
            if (std::fabs(M[r][k]) > std::fabs(M[p][k])) p = r;
        if (!(std::fabs(M[p][k]) > 0.0)) return false;
        if (p != k) {
            for (int c = 0; c < n; c++) std::swap(M[k][c], M[p][c]);
            std::swap(b[k], b[p]);
        }
        for (int r = k + 1; r < n; r++) {
            const double f = M[r][k] / M[k][k];
            if (f == 0.0) continue;
            for (int c = k; c < n; c++) M[r][c] -= f * M[k][c];
            b[r] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; k--) {
        double s = b[k];
        for (int c = k + 1; c < n; c++) s -= M[k][c] * b[c];
        b[k] = s / M[k][k];
    }
    return true;
}

}  // namespace

struct xdem_hht_wheel {
    xdem_hht_wheel_config cfg;
    double q[kQ], qd[kQ];
    double rhs[kT];            // accelerations (7) and Lagrange multipliers (5)
    double pre[kQ], ipp[kQ], ipv[kQ], ee[kT];
    double lhs[kT][kT];
    double cj[kC][kQ];         // constraint Jacobian
    double fc[3], mc[3];       // contact force and moment of the current step (global)
    double t = 0.0, y0 = 0.0, theta0 = 0.0;
    double alpha = 0.0, beta = 0.0, gamma = 0.0;
    double dt2accp = 0.0, dt2accv = 0.0, dt2acc = 0.0, divalpha = 0.0, divbeta = 0.0;

    // xMultiBodySimulation::ConstructMassMatrix(mul): the first kQ rows and columns of lhs
    void mass_matrix(double mul)
    {
        std::memset(lhs, 0, sizeof(lhs));
        lhs[0][0] = lhs[1][1] = lhs[2][2] = mul * cfg.mass;
        const Mat34 G = gmat(q + 3);
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) {
                double s = 0.0;
                for (int k = 0; k < 3; k++) s += G.a[k][r] * (cfg.inertia[k] * G.a[k][c]);
                lhs[3 + r][3 + c] = 4.0 * mul * s;
            }
    }

    // xIntegratorHHT::MassJacobian(mul): d(M(e) e_dd)/de on the rotational block
    void mass_jacobian(double mul)
    {
        const double* e = q + 3;
        const double* ea = rhs + 3;
        const Mat34 G = gmat(e), Ga = gmat(ea);
        double v[3];   // J (G(e) e_dd)
        for (int k = 0; k < 3; k++) {
            double s = 0.0;
            for (int c = 0; c < 4; c++) s += G.a[k][c] * ea[c];
            v[k] = cfg.inertia[k] * s;
        }
        const double Mm[4][4] = {{0.0, -v[0], -v[1], -v[2]},
                                 {v[0], 0.0, v[2], -v[1]},
                                 {v[1], -v[2], 0.0, v[0]},
                                 {v[2], v[1], -v[0], 0.0}};
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) {
                double s = 0.0;
                for (int k = 0; k < 3; k++) s += G.a[k][r] * (cfg.inertia[k] * Ga.a[k][c]);
                lhs[3 + r][3 + c] += 4.0 * mul * (-s + Mm[r][c]);
            }
    }

    // xMultiBodySimulation::ConstructForceVector: applied forces and the quadratic velocity term
    void force_vector(double* v) const
    {
        const double* e = q + 3;
        const double* ed = qd + 3;
        for (int a = 0; a < 3; a++) v[a] = cfg.mass * cfg.gravity[a] + fc[a] + cfg.load[a];
        double A[3][3];
        amat(e, A);
        double nb[3];   // body-frame moment A^T M
        for (int k = 0; k < 3; k++) nb[k] = A[0][k] * mc[0] + A[1][k] * mc[1] + A[2][k] * mc[2];
        const Mat34 G = gmat(e), Gd = gmat(ed);
        double w[3];   // J G(e_dot) e
        for (int k = 0; k < 3; k++) {
            double s = 0.0;
            for (int c = 0; c < 4; c++) s += Gd.a[k][c] * e[c];
            w[k] = cfg.inertia[k] * s;
        }
        for (int c = 0; c < 4; c++) {
            double s = 0.0, s2 = 0.0;
            for (int k = 0; k < 3; k++) {
                s += G.a[k][c] * nb[k];
                s2 += Gd.a[k][c] * w[k];
            }
            v[3 + c] = 2.0 * s + 8.0 * s2;
        }
    }

    // the rig's constraint Jacobian at q (time tt for the drive)
    void jacobian(double tt)
    {
        std::memset(cj, 0, sizeof(cj));
        const double* e = q + 3;
        const double th = 0.5 * (theta0 + cfg.omega * tt);
        cj[0][1] = 1.0;
        cj[1][3] = -2.0 * e[3]; cj[1][4] = 2.0 * e[2]; cj[1][5] = 2.0 * e[1]; cj[1][6] = -2.0 * e[0];
        cj[2][3] = 2.0 * e[1]; cj[2][4] = 2.0 * e[0]; cj[2][5] = 2.0 * e[3]; cj[2][6] = 2.0 * e[2];
        cj[3][3] = -std::sin(th); cj[3][5] = std::cos(th);
        for (int c = 0; c < 4; c++) cj[4][3 + c] = 2.0 * e[c];
    }

    void constraints(double tt, double mul, double* v) const
    {
        const double* e = q + 3;
        const double th = 0.5 * (theta0 + cfg.omega * tt);
        v[0] = mul * (q[1] - y0);
        v[1] = mul * (2.0 * (e[1] * e[2] - e[0] * e[3]));
        v[2] = mul * (2.0 * (e[2] * e[3] + e[0] * e[1]));
        v[3] = mul * (e[2] * std::cos(th) - e[0] * std::sin(th));
        v[4] = mul * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + e[3] * e[3] - 1.0);
    }

    void put_jacobian()
    {
        for (int r = 0; r < kC; r++)
            for (int c = 0; c < kQ; c++) lhs[kQ + r][c] = lhs[c][kQ + r] = cj[r][c];
    }

    // xIntegratorHHT::Initialize: consistent accelerations and multipliers
    int initialize()
    {
        beta = (1.0 - alpha) * (1.0 - alpha) / 4.0;
        gamma = 0.5 - alpha;
        const double dt = cfg.dt;
        dt2accp = dt * dt * (1.0 - 2.0 * beta) * 0.5;
        dt2accv = dt * (1.0 - gamma);
        dt2acc = dt * dt * beta;
        divalpha = 1.0 / (1.0 + alpha);
        divbeta = -1.0 / (beta * dt * dt);
        mass_matrix(1.0);
        jacobian(t);
        std::memset(rhs, 0, sizeof(rhs));
        force_vector(rhs);
        put_jacobian();
        return solve(lhs, rhs, kT) ? XDEM_OK : XDEM_E_INVALID;
    }

    // xIntegratorHHT::PredictionStep + CorrectionStep
    int step(int* iterations)
    {
        const double dt = cfg.dt;
        const double* lam = rhs + kQ;
        force_vector(pre);
        jacobian(t);
        for (int r = 0; r < kC; r++)
            for (int c = 0; c < kQ; c++) pre[c] -= cj[r][c] * lam[r];
        for (int c = 0; c < kQ; c++) pre[c] *= alpha / (1.0 + alpha);
        for (int i = 0; i < kQ; i++) {
            ipp[i] = q[i] + dt * qd[i] + dt2accp * rhs[i];
            ipv[i] = qd[i] + dt2accv * rhs[i];
        }
        for (int i = 0; i < kQ; i++) {
            q[i] = ipp[i] + dt2acc * rhs[i];
            qd[i] = ipv[i] + rhs[i] * dt * gamma;
        }
        const double t1 = t + dt;
        const double eps = cfg.tolerance > 0.0 ? cfg.tolerance : 1e-5;
        const int itmax = cfg.max_iterations > 0 ? cfg.max_iterations : 100;
        int it = 0;
        for (;;) {
            if (++it > itmax) {
                if (iterations) *iterations = it - 1;
                return XDEM_E_DEVICE_ERROR;   // the Newton iteration did not converge
            }
            force_vector(ee);
            jacobian(t1);
            for (int r = 0; r < kC; r++)
                for (int c = 0; c < kQ; c++) ee[c] -= cj[r][c] * lam[r];
            mass_matrix(divalpha);
            for (int r = 0; r < kQ; r++) {
                double s = 0.0;
                for (int c = 0; c < kQ; c++) s += lhs[r][c] * rhs[c];
                ee[r] -= s;
            }
            mass_jacobian(divalpha * beta * dt * dt);
            put_jacobian();
            for (int i = 0; i < kQ; i++) ee[i] -= pre[i];
            constraints(t1, divbeta, ee + kQ);
            double nrm = 0.0;
            for (int i = 0; i < kT; i++) nrm += ee[i] * ee[i];
            nrm = std::sqrt(nrm);
            if (!solve(lhs, ee, kT)) return XDEM_E_INVALID;
            for (int i = 0; i < kT; i++) rhs[i] += ee[i];
            for (int i = 0; i < kQ; i++) {
                q[i] = ipp[i] + dt2acc * rhs[i];
                qd[i] = ipv[i] + dt * gamma * rhs[i];
            }
            if (nrm <= eps) break;
        }
        t = t1;
        if (iterations) *iterations = it;
        return XDEM_OK;
    }

    void pose(double* d) const
    {
        d[XDEM_MM] = cfg.mass;
        for (int a = 0; a < 3; a++) { d[XDEM_PX + a] = q[a]; d[XDEM_VX + a] = qd[a]; }
        for (int c = 0; c < 4; c++) { d[XDEM_E0 + c] = q[3 + c]; d[XDEM_ED0 + c] = qd[3 + c]; }
    }
};

extern "C" {

int xdem_hht_wheel_create(const xdem_hht_wheel_config* cfg, xdem_hht_wheel** out)
{
    if (!cfg || !out) return XDEM_E_INVALID;
    *out = nullptr;
    if (!(cfg->mass > 0.0) || !(cfg->dt > 0.0) || !(cfg->alpha <= 0.0 && cfg->alpha >= -1.0 / 3.0)) return XDEM_E_INVALID;
    for (int k = 0; k < 3; k++)
        if (!(cfg->inertia[k] > 0.0)) return XDEM_E_INVALID;
    xdem_hht_wheel* w = new (std::nothrow) xdem_hht_wheel();
    if (!w) return XDEM_E_INVALID;
    w->cfg = *cfg;
    for (int a = 0; a < 3; a++) { w->q[a] = cfg->dbi[XDEM_PX + a]; w->qd[a] = cfg->dbi[XDEM_VX + a]; }
    for (int c = 0; c < 4; c++) { w->q[3 + c] = cfg->dbi[XDEM_E0 + c]; w->qd[3 + c] = cfg->dbi[XDEM_ED0 + c]; }
    w->y0 = w->q[1];
    w->theta0 = 2.0 * std::atan2(w->q[5], w->q[3]);   // e = (cos(theta/2), 0, sin(theta/2), 0)
    std::memset(w->fc, 0, sizeof(w->fc));
    std::memset(w->mc, 0, sizeof(w->mc));
    w->alpha = cfg->alpha;
    const int rc = w->initialize();
    if (rc != XDEM_OK) {
        delete w;
        return rc;
    }
    *out = w;
    return XDEM_OK;
}

int xdem_hht_wheel_destroy(xdem_hht_wheel* w)
{
    delete w;
    return XDEM_OK;
}

int xdem_hht_wheel_step(xdem_hht_wheel* w, const double wrench[6], double dbi_out[15], int* iterations)
{
    if (!w) return XDEM_E_INVALID;
    for (int a = 0; a < 3; a++) {   // zeroed every step, then this step's DEM wrench
        w->fc[a] = wrench ? wrench[a] : 0.0;
        w->mc[a] = wrench ? wrench[3 + a] : 0.0;
    }
    const int rc = w->step(iterations);
    if (dbi_out) w->pose(dbi_out);
    return rc;
}

int xdem_hht_wheel_state(const xdem_hht_wheel* w, double dbi_out[15], double qdd_out[7], double lambda_out[5],
                         double* time)
{
    if (!w) return XDEM_E_INVALID;
    if (dbi_out) w->pose(dbi_out);
    if (qdd_out)
        for (int i = 0; i < kQ; i++) qdd_out[i] = w->rhs[i];
    if (lambda_out)
        for (int i = 0; i < kC; i++) lambda_out[i] = w->rhs[kQ + i];
    if (time) *time = w->t;
    return XDEM_OK;
}

}  // extern "C"
