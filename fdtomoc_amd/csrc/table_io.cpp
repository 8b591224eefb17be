// table_io.cpp -- travel-time table files into resident tables, read the way
// FDtomoC's consumers read them (include/fdtomoc_amd.h, fdgpu_table_load):
// sphrayderv.c:1184-1217 and sphfdloc's read_timefiles take the first
// whitespace-delimited token of the 232-byte header; "HEAD" means the floats
// follow the header, anything else means the file has no header and the
// floats start at byte 0.  nx*ny*nz native floats, x fastest, are read.
#include <hip/hip_runtime.h>
#include <ctype.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "fdgpu_internal.h"

using namespace fdgpu_detail;

namespace {

constexpr size_t kHeadBytes = 232;   // struct vhead (vhead.h:4-11)

// The first token of the header as sscanf("%4s") reads it: leading white
// space skipped, then at most four characters up to white space or the end
// of the string.
bool header_says_head(const unsigned char *h, size_t n) {
    size_t q = 0;
    while (q < n && h[q] && isspace(h[q])) q++;
    char tok[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < 4 && q < n && h[q] && !isspace(h[q]); k++, q++) tok[k] = (char)h[q];
    return strcmp(tok, "HEAD") == 0;
}

}  // namespace

extern "C" int fdgpu_table_load(fdgpu_ctx *c, const char *path, int nx, int ny, int nz, int *table) {
    if (!c || !path || !table || nx < 1 || ny < 1 || nz < 1) return FDGPU_EINPUT;
    const long long n = (long long)nx * ny * nz;
    if (n >= (1LL << 31)) return FDGPU_EINPUT;
    std::vector<float> host((size_t)n);
    FILE *f = fopen(path, "rb");
    if (!f) return FDGPU_EIO;
    unsigned char head[kHeadBytes];
    const size_t got = fread(head, 1, kHeadBytes, f);
    if (!(got == kHeadBytes && header_says_head(head, kHeadBytes))) rewind(f);
    const size_t nread = fread(host.data(), sizeof(float), (size_t)n, f);
    fclose(f);
    if (nread != (size_t)n) return FDGPU_EIO;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;   // no allocation while the request loop runs
    if (hipSetDevice(c->device) != hipSuccess) return FDGPU_EHIP;
    Table T;
    fdgpu_par_defaults(&T.par);
    T.par.nx = nx;
    T.par.ny = ny;
    T.par.nz = nz;
    T.n = n;
    T.from_file = true;
    int rc = dev_alloc_resident(c, (void **)&T.d_time, sizeof(float) * (size_t)n);
    if (rc) return rc;
    if (hipMemcpyAsync(T.d_time, host.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        dev_free(c, T.d_time, sizeof(float) * (size_t)n);
        return FDGPU_EHIP;
    }
    const int id = c->next_id++;
    c->tables[id] = T;
    *table = id;
    return FDGPU_OK;
}
