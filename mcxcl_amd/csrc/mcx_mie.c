/*
 * mcx_mie.c - Mie scattering of a homogeneous sphere, for polarised photon
 * transport: the scattering efficiency, the anisotropy and the Mueller-matrix
 * elements S11, S12, S33, S34 on a table of scattering angles.
 *
 * The series is the Lorenz-Mie solution as published by Bohren and Huffman
 * ("Absorption and Scattering of Light by Small Particles", 1983, ch. 4 and
 * appendix A): Riccati-Bessel functions psi_n, chi_n by upward recurrence, the
 * logarithmic derivative D_n(mx) by downward recurrence from well above the
 * last order (stable for every real m), the angular functions pi_n, tau_n by
 * their recurrences, and Wiscombe's order count nstop = x + 4.05 x^(1/3) + 2
 * (Appl. Opt. 19, 1505, 1980).  xi_n = psi_n + i chi_n with chi_n = -x y_n(x):
 * the time convention of the tables the reference computes (its S34 =
 * Im(conj(S1) S2) has this sign, src/mcx_mie.cpp:123-124, :197-201).  The
 * reference switches to a small-particle expansion below |m| x = 0.1; the
 * series here is used at every size.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>

#include "mcx_utils.h"

/* the table for one sphere: x = 2 pi r n_med / lambda, m = n_sph / n_med (real),
   mu[k] = cos(theta_k) for nangle angles; smatrix gets {S11, S12, S33, S34} per
   angle as floats.  Returns 0, or -1 for x outside (0, 20000]. */
int mcx_mie(double x, double m, int nangle, const double* mu, float* smatrix, double* qsca, double* g) {
    if (!(x > 0.0) || x > 20000.0 || nangle <= 0) {
        return -1;
    }

    const long nstop = (long)floor(x + 4.05 * pow(x, 1.0 / 3.0) + 2.0);
    const double mx = m * x;
    const long nmx = (long)((nstop > fabs(mx) ? nstop : fabs(mx)) + 16);
    double* D = (double*)calloc((size_t)nmx + 1, sizeof(double));
    double complex* s1 = (double complex*)calloc((size_t)nangle, sizeof(double complex));
    double complex* s2 = (double complex*)calloc((size_t)nangle, sizeof(double complex));
    double* pim1 = (double*)calloc((size_t)nangle, sizeof(double));   /* pi_{n-1} */
    double* pin = (double*)calloc((size_t)nangle, sizeof(double));    /* pi_n */

    if (!D || !s1 || !s2 || !pim1 || !pin) {
        free(D);
        free(s1);
        free(s2);
        free(pim1);
        free(pin);
        return -2;
    }

    /* D_{n-1} = n/(mx) - 1/(D_n + n/(mx)), started at zero far above nstop */
    for (long n = nmx; n >= 1; n--) {
        const double r = (double)n / mx;
        D[n - 1] = r - 1.0 / (D[n] + r);
    }

    for (int k = 0; k < nangle; k++) {
        pin[k] = 1.0;     /* pi_1 */
        pim1[k] = 0.0;    /* pi_0 */
    }

    /* psi_{-1} = cos x, psi_0 = sin x; chi_{-1} = -sin x, chi_0 = cos x */
    double psim1 = cos(x), psi0 = sin(x), chim1 = -sin(x), chi0 = cos(x);
    double complex aprev = 0.0, bprev = 0.0;
    double qs = 0.0, gs = 0.0;

    for (long n = 1; n <= nstop; n++) {
        const double fn = (double)n;
        const double psi = (2.0 * fn - 1.0) / x * psi0 - psim1;
        const double chi = (2.0 * fn - 1.0) / x * chi0 - chim1;
        const double complex xi = psi + I * chi, xim1 = psi0 + I * chi0;
        const double da = D[n] / m + fn / x, db = D[n] * m + fn / x;
        const double complex an = (da * psi - psi0) / (da * xi - xim1);
        const double complex bn = (db * psi - psi0) / (db * xi - xim1);

        qs += (2.0 * fn + 1.0) * (creal(an) * creal(an) + cimag(an) * cimag(an) + creal(bn) * creal(bn) +
                                  cimag(bn) * cimag(bn));

        if (n > 1) {
            /* (n-1)(n+1)/n Re(a_{n-1} a_n* + b_{n-1} b_n*) */
            gs += (fn - 1.0) * (fn + 1.0) / fn * creal(aprev * conj(an) + bprev * conj(bn));
        }

        gs += (2.0 * fn + 1.0) / (fn * (fn + 1.0)) * creal(an * conj(bn));

        const double c = (2.0 * fn + 1.0) / (fn * (fn + 1.0));

        for (int k = 0; k < nangle; k++) {
            const double tau = fn * mu[k] * pin[k] - (fn + 1.0) * pim1[k];
            s1[k] += c * (an * pin[k] + bn * tau);
            s2[k] += c * (an * tau + bn * pin[k]);
            /* pi_{n+1} = ((2n+1) mu pi_n - (n+1) pi_{n-1}) / n */
            const double next = ((2.0 * fn + 1.0) * mu[k] * pin[k] - (fn + 1.0) * pim1[k]) / fn;
            pim1[k] = pin[k];
            pin[k] = next;
        }

        psim1 = psi0;
        psi0 = psi;
        chim1 = chi0;
        chi0 = chi;
        aprev = an;
        bprev = bn;
    }

    *qsca = 2.0 / (x * x) * qs;
    *g = 4.0 / (x * x * (*qsca)) * gs;

    for (int k = 0; k < nangle; k++) {
        const double a1 = creal(s1[k]) * creal(s1[k]) + cimag(s1[k]) * cimag(s1[k]);
        const double a2 = creal(s2[k]) * creal(s2[k]) + cimag(s2[k]) * cimag(s2[k]);
        const double complex c12 = conj(s1[k]) * s2[k];
        smatrix[4 * k] = (float)(0.5 * (a2 + a1));
        smatrix[4 * k + 1] = (float)(0.5 * (a2 - a1));
        smatrix[4 * k + 2] = (float)creal(c12);
        smatrix[4 * k + 3] = (float)cimag(c12);
    }

    free(D);
    free(s1);
    free(s2);
    free(pim1);
    free(pin);
    return 0;
}
