// sphfd_io.cpp -- see sphfd_io.h.
#include "sphfd_io.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace fdio {

namespace {
constexpr int kMaxLine = 1024;     // sphfd.c:6216
constexpr int kMaxParLevel = 4;    // sphfd.c:6223
constexpr int kSpecLine = 132;     // parseprogs.c:4

void swap4(void *p) {
    unsigned char *c = (unsigned char *)p, t;
    t = c[0]; c[0] = c[3]; c[3] = t;
    t = c[1]; c[1] = c[2]; c[2] = t;
}
void swap8(void *p) {
    unsigned char *c = (unsigned char *)p;
    for (int i = 0; i < 4; i++) {
        const unsigned char t = c[i];
        c[i] = c[7 - i];
        c[7 - i] = t;
    }
}
}  // namespace

// ------------------------------------------------------------------ getpar

void GetPar::add(const std::string &n, const std::string &v) { list_.emplace_back(n, v); }

int GetPar::do_par_file(const std::string &fname, int level) {
    // sphfd.c:6716-6767
    if (level > kMaxParLevel) return FDGPU_EPAR;
    FILE *f = fopen(fname.c_str(), "r");
    if (!f) return FDGPU_EIO;
    char line[kMaxLine];
    int rc = 0;
    while (rc == 0 && fgets(line, kMaxLine, f)) {
        const char *pl = line;
        while (true) {
            while (*pl == ' ' || *pl == '\t') pl++;
            if (*pl == '\0' || *pl == '\n' || *pl == '#') break;
            std::string name, value;
            while (*pl != '=' && *pl != '\0' && *pl != ' ' && *pl != '\t') name += *pl++;
            if (*pl == '=') pl++;
            char t1, t2;
            if (*pl == '"' || *pl == '\'') {
                t1 = t2 = *pl++;
            } else {
                t1 = ' ';
                t2 = '\t';
            }
            while (*pl != t1 && *pl != t2 && *pl != '\0' && *pl != '\n') value += *pl++;
            if (*pl == '"' || *pl == '\'') pl++;
            add(name, value);
            if (name == "par") {
                rc = do_par_file(value, level + 1);
                if (rc) break;
            }
        }
    }
    fclose(f);
    return rc;
}

int GetPar::setpar(int ac, char **av) {
    // sphfd.c:6267-6361
    for (int a = 1; a < ac; a++) {
        const char *pl = av[a];
        while (*pl == ' ' || *pl == '\t') pl++;
        std::string name, value;
        while (*pl != '=' && *pl != '\0') name += *pl++;
        if (*pl == '=') pl++;
        if (*pl == '"' || *pl == '\'') {
            const char t = *pl++;
            while (*pl != '\0') {
                if (*pl == t) {
                    if (pl[-1] != '\\') break;
                    value.back() = t;   // an escaped quote replaces its backslash
                    pl++;
                } else {
                    value += *pl++;
                }
            }
        } else {
            value = pl;
        }
        if (!name.empty() && name[0] == '-')
            add("SWITCH", name.substr(1));
        else
            add(name, value);
        if (name == "par") {
            const int rc = do_par_file(value, 1);
            if (rc) return rc;
        }
    }
    std::string v;
    stop_ = getpar("STOP", "s", nullptr);
    return 0;
}

bool GetPar::getpar(const char *name, const char *type, void *val) const {
    for (auto it = list_.rbegin(); it != list_.rend(); ++it) {
        if (it->first != name) continue;
        if (!val) return true;
        const char *str = it->second.c_str();
        switch (type[0]) {
        case 'd': *(int *)val = atoi(str); break;
        case 'f': *(float *)val = (float)atof(str); break;
        case 'F': *(double *)val = atof(str); break;
        case 's': *(std::string *)val = it->second; break;
        default: return false;
        }
        return true;
    }
    return false;
}

int read_params(const GetPar &gp, fdprep::ParValues &p, std::string &missing) {
    // sphfd.c:383-531
    auto must = [&](const char *n, const char *t, void *v) {
        if (!missing.empty()) return;
        if (!gp.getpar(n, t, v)) missing = n;
    };
    must("nx", "d", &p.nx);
    must("ny", "d", &p.ny);
    must("nz", "d", &p.nz);
    must("h", "F", &p.h);
    if (!missing.empty()) return FDGPU_EPAR;
    gp.getpar("dq", "F", &p.dq);
    gp.getpar("df", "F", &p.df);
    gp.getpar("iminus", "d", &p.iminus);
    gp.getpar("iplus", "d", &p.iplus);
    gp.getpar("jminus", "d", &p.jminus);
    gp.getpar("jplus", "d", &p.jplus);
    gp.getpar("kminus", "d", &p.kminus);
    gp.getpar("kplus", "d", &p.kplus);
    gp.getpar("floatsrc", "d", &p.floatsrc);
    gp.getpar("srctype", "d", &p.srctype);
    gp.getpar("NCUBE", "d", &p.ncube);
    gp.getpar("reverse", "d", &p.reverse);
    gp.getpar("headpref", "d", &p.headpref);
    gp.getpar("maxoff", "f", &p.maxoff);
    gp.getpar("x0", "F", &p.x0);
    gp.getpar("y0", "F", &p.y0);
    gp.getpar("z0", "F", &p.z0);
    gp.getpar("parse", "d", &p.parse);
    gp.getpar("fill", "d", &p.fill);
    gp.getpar("invh", "d", &p.invh);
    gp.getpar("invf", "d", &p.invf);
    gp.getpar("invq", "d", &p.invq);
    gp.getpar("invx0", "d", &p.invx0);
    gp.getpar("invy0", "d", &p.invy0);
    gp.getpar("invz0", "d", &p.invz0);
    gp.getpar("swab", "d", &p.swab);
    if (p.srctype == 1) {
        if (p.floatsrc == 0) {
            must("xs", "d", &p.xs);
            must("ys", "d", &p.ys);
            must("zs", "d", &p.zs);
        } else {
            must("fxs", "F", &p.fxs);
            must("fys", "F", &p.fys);
            must("fzs", "F", &p.fzs);
        }
    } else if (p.srctype == 2) {
        must("srcwall", "d", &p.srcwall);
        must("wallfile", "s", &p.wallfile);
    } else if (p.srctype == 3) {
        must("srcwall", "d", &p.srcwall);
        must("oldtfile", "s", &p.oldtfile);
    } else {
        return FDGPU_ESRCTYPE;
    }
    if (p.fill == 1) must("boxfile", "s", &p.boxfile);
    must("timefile", "s", &p.timefile);
    must("velfile", "s", &p.velfile);
    if (!missing.empty()) return FDGPU_EPAR;
    if (gp.stop()) return FDGPU_ESTOP;
    return 0;
}

// ---------------------------------------------------------------- get_vars

bool get_vars(const char *path, const char *vname, std::string &value) {
    // parseprogs.c:9-77: first line whose first field equals vname (up to its
    // first blank); the value is the second field (quotes allowed)
    FILE *f = fopen(path, "r");
    if (!f) return false;
    std::string want(vname);
    const size_t sp = want.find(' ');
    if (sp != std::string::npos) want.resize(sp);
    char line[kSpecLine];
    bool found = false;
    auto field = [](const char *&pl, std::string &out) -> bool {
        out.clear();
        while (*pl == ' ' || *pl == '\t') pl++;
        if (*pl == '\0') return false;
        if (*pl == '"') {
            pl++;
            while (*pl && *pl != '"') out += *pl++;
            if (*pl != '"') return false;
            pl++;
            return true;
        }
        while (*pl && *pl != ' ' && *pl != '\t') out += *pl++;
        return true;
    };
    while (!found && fgets(line, kSpecLine, f)) {
        const size_t L = strlen(line);
        if (L > 0) line[L - 1] = '\0';   // get_line drops the last character (the newline)
        if (line[0] == '#' || line[0] == '\0') continue;
        const char *pl = line;
        std::string name, val;
        if (!field(pl, name) || name != want) continue;
        if (!field(pl, val)) continue;
        value = val;
        found = true;
    }
    fclose(f);
    return found;
}

// ------------------------------------------------------------ velocity file

VHead table_header(const fdprep::ParValues &p, const fdprep::Prepared &g, const VHead &file_head, bool has_head) {
    VHead head = file_head;
    if (!has_head) {
        // sphfd.c:615-656: the synthesized header of a headerless file
        head.x0 = g.x0_deg;
        head.y0 = g.y0_deg;
        head.z0 = g.z0_km;
        head.dx = g.df;
        head.dy = g.dq;
        head.dz = g.h;
        head.nx = p.nx;
        head.ny = p.ny;
        head.nz = p.nz;
        head.az = 0.f;
        head.clat = g.y0_deg;
        head.clon = g.x0_deg;
        head.cz = g.z0_km;
        // The reference blanks bytes 0..143 with strcpy(" "), whose last
        // terminator lands on byte 144, the low byte of clat.
        char *raw = (char *)&head;
        for (int i = 0; i < 144; i++) raw[i] = ' ';
        raw[144] = '\0';
        const char *txt = "HEADFINESPHRVPMDNOFLsphfd";
        memcpy(head.header, txt, strlen(txt) + 1);
    }
    if (p.swab == 1 || p.swab == 3) {
        // sphfd.c:668-682
        swap4(&head.az); swap4(&head.nx); swap4(&head.ny); swap4(&head.nz);
        swap8(&head.clat); swap8(&head.clon); swap8(&head.cz);
        swap8(&head.x0); swap8(&head.y0); swap8(&head.z0);
        swap8(&head.dx); swap8(&head.dy); swap8(&head.dz);
    }
    return head;
}

int read_velocity(const fdprep::ParValues &p, const std::string &path, std::vector<float> &vel, VHead &file_head,
                  bool &has_head) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return FDGPU_EIO;
    memset(&file_head, 0, sizeof file_head);
    const size_t got = fread(&file_head, 1, sizeof file_head, f);
    has_head = got == sizeof file_head && strncmp(file_head.header, "HEAD", 4) == 0;
    if (!has_head) fseek(f, 0, SEEK_SET);   // headerless: velocities start at byte 0
    const long long n = (long long)p.nx * p.ny * p.nz;
    vel.resize((size_t)n);
    const size_t nr = fread(vel.data(), sizeof(float), (size_t)n, f);
    fclose(f);
    if ((long long)nr != n) return FDGPU_EIO;
    if (p.swab == 1 || p.swab == 3)
        for (auto &v : vel) swap4(&v);   // sphfd.c:683-684
    return 0;
}

// ----------------------------------------------------------------- tables

int write_table(const char *path, const fdgpu_par &par, const fdprep::Prepared &g, const void *vel_header,
                const float *time) {
    const int nx = par.nx, ny = par.ny, nz = par.nz, parse = par.parse;
    const long long nxy = (long long)nx * ny;
    std::vector<float> out;
    int nxp = nx, nyp = ny, nzp = nz;
    if (parse > 1) {
        // sphfd.c:5706-5741
        nzp = 0;
        for (int k = 0; k < nz; k += parse) {
            nzp++;
            for (int j = 0; j < ny; j += parse)
                for (int i = 0; i < nx; i += parse) out.push_back(time[i + (long long)j * nx + k * nxy]);
        }
        nyp = 0;
        for (int j = 0; j < ny; j += parse) nyp++;
        nxp = 0;
        for (int i = 0; i < nx; i += parse) nxp++;
    } else {
        out.assign(time, time + nxy * nz);
    }
    VHead in, ho;
    memcpy(&in, vel_header, sizeof in);
    memset(&ho, 0, sizeof ho);
    // sphfd.c:5744-5769: header text with quantity PTIM
    strncpy(ho.header, in.header, 120);
    memcpy(ho.header + 12, "PTIM", 4);
    memcpy(ho.header + 16, in.header + 16, 104);
    ho.clat = in.clat; ho.clon = in.clon; ho.cz = in.cz; ho.az = in.az;
    ho.x0 = in.x0; ho.y0 = in.y0; ho.z0 = in.z0;
    ho.dx = in.dx * parse; ho.dy = in.dy * parse; ho.dz = in.dz * parse;
    ho.nx = nxp; ho.ny = nyp; ho.nz = nzp;
    ho.fxs = g.fxss; ho.fys = g.fyss; ho.fzs = g.fzss;
    if (par.swab == 2 || par.swab == 3) {
        // sphfd.c:5774-5795
        swap4(&ho.az); swap4(&ho.nx); swap4(&ho.ny); swap4(&ho.nz);
        swap8(&ho.clat); swap8(&ho.clon); swap8(&ho.cz);
        swap8(&ho.x0); swap8(&ho.y0); swap8(&ho.z0);
        swap8(&ho.dx); swap8(&ho.dy); swap8(&ho.dz);
        swap8(&ho.fxs); swap8(&ho.fys); swap8(&ho.fzs);
        for (auto &v : out) swap4(&v);
    }
    FILE *f = fopen(path, "wb");
    if (!f) return FDGPU_EIO;
    bool ok = fwrite(&ho, 1, sizeof ho, f) == sizeof ho;
    ok = ok && fwrite(out.data(), sizeof(float), out.size(), f) == out.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : FDGPU_EIO;
}

fdgpu_par to_par(const fdprep::ParValues &v) {
    fdgpu_par p;
    memset(&p, 0, sizeof p);
    p.nx = v.nx; p.ny = v.ny; p.nz = v.nz;
    p.h = v.h; p.dq = v.dq; p.df = v.df; p.x0 = v.x0; p.y0 = v.y0; p.z0 = v.z0;
    p.iminus = v.iminus; p.iplus = v.iplus; p.jminus = v.jminus; p.jplus = v.jplus; p.kminus = v.kminus; p.kplus = v.kplus;
    p.floatsrc = v.floatsrc; p.srctype = v.srctype; p.srcwall = v.srcwall;
    p.fxs = v.fxs; p.fys = v.fys; p.fzs = v.fzs; p.xs = v.xs; p.ys = v.ys; p.zs = v.zs;
    p.ncube = v.ncube; p.reverse = v.reverse; p.headpref = v.headpref; p.parse = v.parse; p.fill = v.fill;
    p.maxoff = v.maxoff;
    p.invh = v.invh; p.invq = v.invq; p.invf = v.invf; p.invx0 = v.invx0; p.invy0 = v.invy0; p.invz0 = v.invz0;
    p.swab = v.swab;
    return p;
}

}  // namespace fdio
