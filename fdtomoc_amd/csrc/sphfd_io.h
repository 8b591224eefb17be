// sphfd_io.h -- the file-level contract of the reference `sphfd`:
// par files (Clayton getpar, src/sphfd.c:6186-6843), the spec file
// (get_vars, lib/common/parseprogs.c:9-205), velocity files and travel-time
// tables (include/common/vhead.h:4-11, src/sphfd.c:611-690, 5702-5803).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fdtomoc_amd.h"
#include "table_prep.h"

namespace fdio {

#pragma pack(push, 1)
struct VHead {
    char header[120];
    double fxs, fys, fzs;
    double clat, clon, cz;
    double x0, y0, z0, dx, dy, dz;
    float az;
    int32_t nx, ny, nz;
};
#pragma pack(pop)
static_assert(sizeof(VHead) == 232, "vhead layout");

// getpar: name=value entries from argv-style tokens and par files, looked up
// last definition first (sphfd.c:6565).
class GetPar {
  public:
    // sphfd.c:6267-6361.  Returns 0 or FDGPU_EPAR / FDGPU_EIO.
    int setpar(int ac, char **av);
    // Type codes as the reference: "d" int (atoi), "f" float, "F" double, "s" string.
    bool getpar(const char *name, const char *type, void *val) const;
    bool stop() const { return stop_; }

  private:
    int do_par_file(const std::string &fname, int level);
    void add(const std::string &n, const std::string &v);
    std::vector<std::pair<std::string, std::string>> list_;
    bool stop_ = false;
};

// The parameter intake of sphfd.c:383-531: required and optional keys into
// ParValues.  Returns 0 or FDGPU_EPAR / FDGPU_ESTOP / FDGPU_ESRCTYPE.
int read_params(const GetPar &gp, fdprep::ParValues &p, std::string &missing);

// get_vars: first `name value` line of a spec file (parseprogs.c:9-77).
bool get_vars(const char *path, const char *name, std::string &value);

// Reads a velocity file as sphfd.c:611-690 does: a 232-byte header if the
// file starts with "HEAD", else velocities from byte 0; nx*ny*nz floats,
// swapped for swab 1|3.
int read_velocity(const fdprep::ParValues &p, const std::string &path, std::vector<float> &vel, VHead &file_head,
                  bool &has_head);

// The header a table is written from: the velocity file's header, or the
// reference's synthesized one for a headerless file; swapped for swab 1|3.
VHead table_header(const fdprep::ParValues &p, const fdprep::Prepared &g, const VHead &file_head, bool has_head);

// Writes a table as sphfd.c:5702-5803 does (parse decimation, PTIM header,
// source in the header, swab 2|3).  vel_header = 232 header bytes of the
// velocity file or the reference's synthesized fallback.
int write_table(const char *path, const fdgpu_par &par, const fdprep::Prepared &g, const void *vel_header,
                const float *time);

// fdgpu_par <-> ParValues
fdgpu_par to_par(const fdprep::ParValues &v);

}  // namespace fdio
