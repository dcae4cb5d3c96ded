// Link-time stand-ins for the launchers of the device translation units
// (chorus_fp64.hip, chorus_mixed.hip) in the CPU-only sanitizer build of the
// library's host code.  Without a GPU, chorus_init returns CHORUS_ERR_HIP
// before any of them could run; reaching one aborts the driver.
#include <cstdio>
#include <cstdlib>

#include "chorus_internal.hpp"

namespace chorus {

[[noreturn]] static void unreachable(const char *f) {
    fprintf(stderr, "device launcher %s reached in the CPU-only build\n", f);
    abort();
}

hipError_t launch_stage_ideal_fp64(int, const StageArgs &, const FaceArgs &, hipStream_t) { unreachable(__func__); }
hipError_t launch_trace(int, int, const double *, double *, double *, hipStream_t) { unreachable(__func__); }
hipError_t launch_trace_mixed(int, int, const double *, const MxTraces &, double *, hipStream_t) {
    unreachable(__func__);
}
hipError_t launch_stage_ideal_mixed(int, const StageArgs &, const FaceArgs &, hipStream_t, int) { unreachable(__func__); }
int tile_count() { unreachable(__func__); }
hipError_t launch_ct(int, const CTArgs &, hipStream_t) { unreachable(__func__); }
hipError_t launch_stage_visc_fp64(int, const StageArgs &, const ViscArgs &, const FaceArgs &, hipStream_t) {
    unreachable(__func__);
}
hipError_t launch_stage_visc_mixed(int, const StageArgs &, const ViscArgs &, const FaceArgs &, hipStream_t, int) {
    unreachable(__func__);
}
int ct_blocks(int, int) { unreachable(__func__); }
int stage_blocks(int, int) { unreachable(__func__); }
hipError_t upload_operators_fp64() { unreachable(__func__); }
hipError_t upload_operators_mixed() { unreachable(__func__); }
hipError_t diag_scratch(int, double **) { unreachable(__func__); }
hipError_t launch_diag_reduce(const double *, size_t, int, double *, hipStream_t) { unreachable(__func__); }
hipError_t launch_plot_nodes(const PlotArgs &, hipStream_t) { unreachable(__func__); }
hipError_t launch_restart_image(int, int, const double *, unsigned *, hipStream_t) { unreachable(__func__); }

}  // namespace chorus
