/*
 * orc_collisions.c -- collision cross sections of the oracle
 * (TEST INFRASTRUCTURE ONLY).
 *
 * FP64 restatement of the reference's fits in units of 1e-20 m^2
 * (e_collisions.cpp:205-308, ion_collisions.cpp:170-199), plus the two
 * channels the physical mode adds (Ar excitation from the tabulated levels
 * shipped with the reference, Ar+ isotropic elastic of Phelps 1994), and the
 * null-collision bound.
 */
#include <math.h>

#include "oracle.h"

/* e-Ar electronic excitation levels (physical mode): the six tables of
 * References/PhelpsAr_e_cross_section.pdf, section 1, tables (2)-(7) (data of
 * Chutjian & Cartwright 1981), energy eV / cross section 1e-20 m^2.  Each is
 * linear between rows, 0 below its first row, held at its last row above. */
static const double exc0_E[23] = {11.549, 11.9, 12.0, 12.5, 12.8, 13.0, 13.1, 13.4, 14.0, 15.0, 16.0, 17.0, 20.0, 25.0, 30.0, 35.0, 40.0, 50.0, 60.0, 70.0, 80.0, 100.0, 120.0};
static const double exc0_S[23] = {0.0, 0.0029, 0.0028, 0.0066, 0.009, 0.009, 0.0115, 0.0094, 0.0125, 0.022, 0.0335, 0.04, 0.033, 0.024, 0.017, 0.0115, 0.0073, 0.0036, 0.0022, 0.0016, 0.001, 0.0006, 0.0005};
static const double exc1_E[23] = {11.624, 12.0, 13.0, 14.0, 15.0, 16.0, 17.0, 18.0, 20.0, 22.0, 24.0, 26.0, 28.0, 30.0, 35.0, 40.0, 45.0, 50.0, 60.0, 70.0, 80.0, 90.0, 100.0};
static const double exc1_S[23] = {0.0, 0.004, 0.012, 0.02, 0.028, 0.035, 0.04, 0.043, 0.047, 0.051, 0.052, 0.053, 0.053, 0.052, 0.05, 0.048, 0.047, 0.045, 0.041, 0.038, 0.035, 0.032, 0.031};
static const double exc2_E[23] = {11.723, 11.9, 12.0, 12.5, 12.8, 13.0, 13.1, 13.4, 14.0, 15.0, 16.0, 17.0, 20.0, 25.0, 30.0, 35.0, 40.0, 50.0, 60.0, 70.0, 80.0, 100.0, 120.0};
static const double exc2_S[23] = {0.0, 0.0006, 0.0006, 0.0013, 0.0018, 0.0018, 0.0023, 0.0019, 0.0025, 0.0045, 0.0061, 0.0075, 0.0066, 0.006, 0.0045, 0.0029, 0.0019, 0.0012, 0.0009, 0.0007, 0.0006, 0.0005, 0.0004};
static const double exc3_E[22] = {11.828, 13.0, 14.0, 15.0, 16.0, 17.0, 18.0, 20.0, 22.0, 24.0, 26.0, 28.0, 30.0, 35.0, 40.0, 45.0, 50.0, 60.0, 70.0, 80.0, 90.0, 100.0};
static const double exc3_S[22] = {0.0, 0.029, 0.049, 0.065, 0.076, 0.086, 0.094, 0.103, 0.114, 0.123, 0.132, 0.14, 0.147, 0.16, 0.17, 0.173, 0.175, 0.175, 0.17, 0.164, 0.156, 0.15};
static const double exc4_E[21] = {12.9, 13.0, 13.1, 13.4, 14.0, 15.0, 16.0, 17.0, 18.0, 20.0, 22.0, 25.0, 30.0, 35.0, 40.0, 50.0, 60.0, 70.0, 80.0, 100.0, 120.0};
static const double exc4_S[21] = {0.0, 0.003, 0.009, 0.0315, 0.07, 0.15, 0.24, 0.32, 0.4, 0.5, 0.53, 0.53, 0.52, 0.5, 0.48, 0.43, 0.4, 0.37, 0.33, 0.28, 0.25};
static const double exc5_E[15] = {14.09, 15.0, 16.0, 17.0, 20.0, 25.0, 30.0, 35.0, 40.0, 50.0, 60.0, 70.0, 80.0, 100.0, 120.0};
static const double exc5_S[15] = {0.0, 0.028, 0.054, 0.075, 0.105, 0.145, 0.16, 0.19, 0.21, 0.235, 0.25, 0.26, 0.265, 0.27, 0.26};

static double table_level(const double *e, const double *s, int n, double E) {
    if (E < e[0]) return 0.0;
    if (E >= e[n - 1]) return s[n - 1];
    for (int k = 0; k + 1 < n; k++)
        if (E < e[k + 1]) return s[k] + (s[k + 1] - s[k]) * (E - e[k]) / (e[k + 1] - e[k]);
    return s[n - 1];
}

/* lumped excitation: the sum of the six levels, 0 below the channel threshold */
static double ar_excitation(double E, double thr) {
    if (E < thr) return 0.0;
    return table_level(exc0_E, exc0_S, 23, E) + table_level(exc1_E, exc1_S, 23, E) +
           table_level(exc2_E, exc2_S, 23, E) + table_level(exc3_E, exc3_S, 22, E) +
           table_level(exc4_E, exc4_S, 21, E) + table_level(exc5_E, exc5_S, 15, E);
}

double orc_sigma(int kind, double E, double thr) {
    switch (kind) {
        case ORC_SIG_AR_EL_REF: /* e_collisions.cpp:205-224 */
            if (E < 0.35)
                return 28019.0 * pow(E, 6.0) - 32798.0 * pow(E, 5.0) + 15048.0 * pow(E, 4.0) -
                       34312.0 * pow(E, 3.0) + 4084.0 * pow(E, 2.0) - 249.0 * E + 7.576;
            if (E < 15.0 && E > 0.35) return -0.013 * pow(E, 3.0) + 0.232 * pow(E, 2.0) + 1.066 * E + 0.085;
            if (E > 15.0) return 126.2 * pow(E, -0.69);
            return 0.0;
        case ORC_SIG_AR_EL_PHYS: {
            double s;
            if (E < 0.35)
                s = 28019.0 * pow(E, 6.0) - 32798.0 * pow(E, 5.0) + 15048.0 * pow(E, 4.0) -
                    34312.0 * pow(E, 3.0) + 4084.0 * pow(E, 2.0) - 249.0 * E + 7.576;
            else if (E < 15.0)
                s = -0.013 * pow(E, 3.0) + 0.232 * pow(E, 2.0) + 1.066 * E + 0.085;
            else
                s = 126.2 * pow(E, -0.69);
            return s > 0.0 ? s : 0.0;
        }
        case ORC_SIG_AR_IZ_REF: /* e_collisions.cpp:235-256 */
            if (E < 15.755) return 0.0;
            if (E < 100.0 && E > 15.755)
                return -3.0e-7 * pow(E, 4.0) + 8.0e-5 * pow(E, 3.0) - 0.008 * pow(E, 2.0) + 0.416 * E - 4.758;
            if (E > 100.0) return 130.0 * pow(E, -0.74);
            return 0.0;
        case ORC_SIG_AR_IZ_PHYS: {
            if (E < thr) return 0.0;
            double s = (E < 100.0) ? -3.0e-7 * pow(E, 4.0) + 8.0e-5 * pow(E, 3.0) - 0.008 * pow(E, 2.0) +
                                         0.416 * E - 4.758
                                   : 130.0 * pow(E, -0.74);
            return s > 0.0 ? s : 0.0;
        }
        case ORC_SIG_AR_EXC: /* physical mode only: sum of the tabulated levels */
            return ar_excitation(E, thr);
        case ORC_SIG_H_EL: /* e_collisions.cpp:267-283 */
            if (E < 1.0e-2) return 3.834 * pow(E, 2.0) - 20.82 * E + 40.98;
            if (E < 20.0) return -6.15 * log(E) + 20.91;
            return 102.2 * pow(E, -1.19);
        case ORC_SIG_H_IZ: { /* e_collisions.cpp:293-308 (Janev), bracket/(I E) */
            double I = 13.6, c = 1.0 - I / E;
            if (E < I) return 0.0;
            return (1.0 / (I * E)) * (0.18450 * log(E / I) - 0.032226 * c - 0.034539 * c * c +
                                      1.4003 * pow(c, 3.0) - 2.8115 * pow(c, 4.0) + 2.2986 * pow(c, 5.0));
        }
        case ORC_SIG_AR_CX: /* ion_collisions.cpp:170-176 */
            if (E > 4.0) return (2.0 + 5.5 / sqrt(E)) * 10.0;
            return (-2.95 * sqrt(E) + 10.65) * 10.0;
        case ORC_SIG_AR_ISO: { /* Phelps (1994) isotropic part, physical mode */
            double e = E > 1e-8 ? E : 1e-8, d = 1.0 + e / 3.0;
            return 20.0 / (sqrt(e) * (1.0 + e)) + 30.0 * e / (d * d);
        }
        case ORC_SIG_H_CX: { /* ion_collisions.cpp:185-199, E in keV/amu */
            if (E < 1.0e-4) return 0.0;
            return (3.2345 * log(235.88 / E + 2.3713)) /
                   (1.0 + 0.038371 * E + 3.8068e-6 * pow(E, 3.5) + 1.1832e-10 * pow(E, 5.4));
        }
    }
    return 0.0;
}
