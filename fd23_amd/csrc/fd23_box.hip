// Hybrid box coupling (placeholder until the plane-wave box lands).
#include "fd23.h"
#include <hip/hip_runtime.h>
struct fd23_ctx;
void *&fd23_ctx_box(fd23_ctx *c);
int fd23_box_launch_pre_stress(fd23_ctx *) { return FD23_OK; }
int fd23_box_launch_post_stress(fd23_ctx *) { return FD23_OK; }
int fd23_box_launch_post_velocity(fd23_ctx *) { return FD23_OK; }
void fd23_box_free(fd23_ctx *) {}
extern "C" int fd23_box_configure(fd23_ctx *, int, int, int, int, int, int, double, double, double, double,
                                  double, double, double, double, double) {
    return FD23_EINVAL;
}
