/* sphfd_main.c -- drop-in `sphfd` executable (src/sphfd.c:200-257): reads the
 * spec-file path on stdin, solves every table of its parlist on the GPU and
 * writes them to timedir.  Exit status 0 on success, 100 for a missing par key
 * and 101 for STOP (as the reference's getpar), 1 otherwise. */
#include <stdio.h>

#include "../../include/fdtomoc_amd.h"

int main(void) {
    char spec[133];
    printf("Input the name spec file\n");
    if (scanf("%132s", spec) != 1) {
        fprintf(stderr, "sphfd: no spec file given on stdin\n");
        return 1;
    }
    const int rc = fdgpu_run_spec(spec);
    if (rc == FDGPU_OK) return 0;
    fprintf(stderr, "sphfd: %s\n", fdgpu_strerror(rc));
    return rc == FDGPU_EPAR ? 100 : rc == FDGPU_ESTOP ? 101 : 1;
}
