/*
 * orc_collisions.c -- collision cross sections of the oracle
 * (TEST INFRASTRUCTURE ONLY).
 *
 * FP64 restatement of the reference's fits in units of 1e-20 m^2
 * (e_collisions.cpp:205-308, ion_collisions.cpp:170-199), plus the two
 * channels the physical mode adds (Ar excitation shape, Ar+ isotropic
 * elastic of Phelps 1994), and the null-collision bound.
 */
#include <math.h>

#include "oracle.h"

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
        case ORC_SIG_AR_EXC: /* physical mode only: 3.262 ln(x)/x, x = E/thr */
            return (E > thr) ? 3.262 * log(E / thr) / (E / thr) : 0.0;
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
