/*
 * ref_mglin_shim.cpp -- C entry point over the reference's own multigrid
 * solver (mglin.cpp / mglin.h / nr3.h compiled from /root/reference by
 * oracle/Makefile into oracle/_ref/libref_mglin.so).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests to pin the oracle's multigrid
 * restatement (orc_mglin) bit for bit against the reference.  No reference
 * source is copied into this repository; the library is rebuilt from the
 * reference tree and never committed.
 */
#include "mglin.h"

extern "C" int ref_mglin(double *u, int n, int ncycle, double length) {
    MatDoub m(n, n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) m[i][j] = u[i * n + j];
    try {
        Mglin solve(m, ncycle, length);
    } catch (...) {
        return -1;
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) u[i * n + j] = m[i][j];
    return 0;
}
