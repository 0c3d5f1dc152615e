// dump_io.cpp -- dataDump_<type>.txt restart files.
//
// Format of particle::dump_data / load_particles (particle.cpp:1361-1382,
// 359-401): one particle per line, "x y vx vy vz weight", whitespace
// separated doubles.  The reference writes with the ostream default of 6
// significant digits; the precision is a parameter here (17 round-trips a
// double exactly).  The reference's reader stops silently at the first token
// it cannot parse; this reader rejects such a file with PIC_ERR_FORMAT.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/pic2d.h"

int read_dump_file(const std::string &path, std::vector<double> &vals, std::string &err) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return PIC_ERR_FILE;
    }
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    fclose(f);
    vals.clear();
    const char *p = text.c_str();
    const char *end = p + text.size();
    while (p < end) {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
        if (p >= end) break;
        char *q = nullptr;
        errno = 0;
        double v = strtod(p, &q);
        if (q == p || (q < end && !(*q == ' ' || *q == '\t' || *q == '\n' || *q == '\r'))) {
            err = path + ": unparsable token after " + std::to_string(vals.size()) + " values";
            return PIC_ERR_FORMAT;
        }
        vals.push_back(v);
        p = q;
    }
    if (vals.size() % 6 != 0) {
        err = path + ": " + std::to_string(vals.size()) + " values is not a whole number of "
                                                          "6-value particle records";
        return PIC_ERR_FORMAT;
    }
    return PIC_OK;
}

int write_dump_file(const std::string &path, const std::vector<pic_particle> &p, int precision) {
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return PIC_ERR_FILE;
    std::vector<char> buf(1 << 20);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    for (const pic_particle &r : p)
        fprintf(f, "%.*g %.*g %.*g %.*g %.*g %.*g\n", precision, r.x, precision, r.y, precision,
                r.vx, precision, r.vy, precision, r.vz, precision, r.weight);
    int rc = fclose(f);
    return rc == 0 ? PIC_OK : PIC_ERR_FILE;
}
