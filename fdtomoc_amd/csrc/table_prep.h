// table_prep.h -- host-side preparation of one sphfd table: parameter
// conversion, geometry, source initialisation and starting box, exactly as
// FDtomoC src/sphfd.c does before its sweep (sphfd.c:415-509, 548-552,
// 696-1003).  Trigonometry stays on the host (glibc libm), so the device
// receives the same doubles the reference computes.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace fdprep {

// Par-file values (keys of sphfd.c:383-413, 460-530), before conversion.
struct ParValues {
    int nx = 0, ny = 0, nz = 0;
    double h = 0, dq = 0, df = 0, x0 = 0, y0 = 90, z0 = 0;
    int iminus = 1, iplus = 1, jminus = 1, jplus = 1, kminus = 1, kplus = 1;
    int floatsrc = 1, srctype = 1, srcwall = 0;
    double fxs = 0, fys = 0, fzs = 0;
    int xs = 0, ys = 0, zs = 0;
    int ncube = 2, reverse = 1, headpref = 6, parse = 1, fill = 0;
    float maxoff = -1.f;
    int invh = 0, invq = 0, invf = 0, invx0 = 0, invy0 = 0, invz0 = 0;
    int swab = 0;
    std::string timefile, velfile, wallfile, oldtfile, boxfile;
};

// Everything the device sweep needs besides the grids.
struct Prepared {
    double h = 0, dq = 0, df = 0, x0 = 0, y0 = 0, z0r = 0;   // radians, colatitude, km
    double x0_deg = 0, y0_deg = 0, z0_km = 0;               // par values after inversion (header fallback)
    int xs = 0, ys = 0, zs = 0;
    double fxss = 0, fyss = 0, fzss = 0;                    // header source (lon rad, colat rad, radius km)
    int maxrad = 99999999;
    std::vector<double> sinq;        // sin(qc(j)) for j = -1..ny at [j + 1]
    // source initialisation
    int init_mode = 0;               // 0 cube distances (times = s000 * dist), 1 preset grid, 2 explicit values
    long long src_index = 0;         // node index of the source (s000 = slowness there)
    std::vector<int> init_idx;
    std::vector<double> init_dist;
    std::vector<float> init_val;
    int box[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    int dflag[6] = {1, 1, 1, 1, 1, 1};
    int grow[6] = {1, 1, 1, 1, 1, 1};   // kminus kplus iplus iminus jminus jplus
    int reverse = 1, headpref = 6, radius0 = 0, srcwall = 0;
    bool near_edge = false;          // sphfd.c:540-545 warning condition
};

// Error codes shared with the C ABI.
enum {
    kOk = 0,
    kErrSourceOutside = -3,   // sphfd.c:533-538
    kErrSrctype = -4,         // sphfd.c:516-519
    kErrInput = -5,           // missing or short input data
};

// Conversion and geometry, sphfd.c:415-509 and 548-552.
int prepare_geometry(const ParValues &p, Prepared &out);

// Source initialisation and starting box (sphfd.c:696-975) and the preferred
// restart wall (sphfd.c:977-1003).  For srctype 2 `preset` must hold the six
// wall planes as the reference reads them, for srctype 3 the first nx*ny*nz
// floats of the old table file; with fill=1 `boxvals` holds the box values.
// For srctype 2/3 the initial grid is written to `grid` (nx*ny*nz floats).
int prepare_source(const ParValues &p, Prepared &out, const float *preset, const float *boxvals, float *grid);

// Whether reverse-pass nodes with unchanged inputs may be skipped without
// changing head-wave counts.  A skipped node reproduces its stored float; the
// reference would still "store" it when the double result lies below that
// float, and would count a head wave if fhead = (improvement)/(d * slowness)
// exceeded 1e-3 (sphfd.c:1756-1759).  The improvement is then below one float
// ulp of the time, so skipping is exact whenever ulp(t_max)/(d_min * s_min) is
// far below 1e-3; this bound is checked with generous margins.
bool skip_safe(const ParValues &p, const Prepared &g, double vmin, double vmax);

// Largest face area of a grid (the reference's sort buffer, sphfd.c:595-604).
long long max_face(int nx, int ny, int nz);

}  // namespace fdprep
