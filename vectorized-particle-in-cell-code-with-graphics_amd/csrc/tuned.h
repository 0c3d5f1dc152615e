// tuned.h -- the library's launch-shape choices per grid size class, as the
// autotuner (autotune.py) measured them; regenerate with it, do not edit.
// A row applies to grids of up to max_nodes nodes per side (the last row to
// every larger grid):
//   sort_c[s]  passes per counting sort of species s (0 electrons, 1 ions)
//              = round(sort_c / d), d = sigma dt / dx the thermal
//              displacement per step in cells, clamped to 1..16
//   sort_key   counting-sort key (tile) width: cell key >> sort_key
//   mg_sweeps  red-black sweeps per leg of the tolerance solve's V-cycle
#pragma once

namespace pic_host {

struct TunedRow {
    int max_nodes;
    double sort_c[2];
    int sort_key;
    int mg_sweeps;
};

// measured by autotune.py on an MI355X (gfx950): argon discharges, 100 particles per cell per species, best of 2 x 20 steps per choice
static const TunedRow TUNED[] = {
    {65, {3.404, 0.8}, 6, 2},
    {129, {6.809, 0.8}, 5, 2},
    {257, {6.809, 0.8}, 6, 2},
    {513, {6.809, 0.8}, 6, 3},
    {1025, {10.213, 0.8}, 6, 2}
};

inline const TunedRow &tuned_row(int nodes) {
    const int n = (int)(sizeof(TUNED) / sizeof(TUNED[0]));
    for (int k = 0; k < n; k++)
        if (nodes <= TUNED[k].max_nodes) return TUNED[k];
    return TUNED[n - 1];
}

}  // namespace pic_host
