// Host FP64 model set-up for branch-site model A (internal header).
#pragma once

namespace bsa {

constexpr int kS = 61;   // sense codons of the universal code
constexpr int kN = 64;   // padded state count used on the device

struct CodonTables {
    int from61[kS];
    int from64[64];
    unsigned char kind[kS][kS];   // 0, or 4 | transition(1) | nonsynonymous(2)
    CodonTables();
};
const CodonTables &tables();

struct ModelA {
    double kappa, p0, p1, w0, w2;
    double freqK[4];
    double wbar[2];       // average omega of background / foreground branches
    double mr[2];         // mean rates at wbar
    double Qfactor[2];    // 1 / mr
    double root[3][kS];   // eigenvalues for omega = w0, 1, w2
    double U[3][kS * kS];
    double V[3][kS * kS];
};

double build_q(double kappa, double omega, const double *pi, double *Q);
double mean_rate(double kappa, double omega, const double *pi);
bool eigen_rev(const double *Q, const double *pi, double *root, double *U, double *V);
// threads: the three eigen systems on three threads (off where the caller
// already runs one gene per thread)
bool setup_model_a(double kappa, double p0, double p1, double w0, double w2, const double *pi,
                   ModelA *m, bool need_eigen, bool threads = true);

// Derivatives of lnL with respect to the natural model parameters, given
// the adjoint outputs: Z[3][64*64] (per eigen set, Z = V^T W U^T, so that
// d lnL / d theta|_Q = <Z, dQ/dtheta>), dQf[2] = d lnL / d Qfactor for
// background / foreground branches, Sf[4] = d lnL / d p_c (class
// proportions).  Qfactor and the proportions are differentiated through
// p0, p1, w0, w2 and kappa here.
struct ModelGrad {
    double kappa, p0, p1, w0, w2;
};
void model_a_param_grad(const ModelA &m, const double *pi, const double *Z, const double *dQf,
                        const double *Sf, ModelGrad *g);

// site class -> eigen set for a branch of type `label` (Set_UVR_BranchSite,
// codeml.c:2555-2558): background w0,1,w0,1; foreground w0,1,w2,w2
inline int uvr_set(int iclass, int label)
{
    return label == 0 ? (iclass & 1) : (iclass <= 1 ? iclass : 2);
}

}  // namespace bsa
