#pragma once
// Runtime settings, same members and defaults as src/Settings.hpp:5-54 of the reference.

namespace Repulsor
{
    enum class OperatorType : int
    {
        FractionalOnly = 0,
        HighOrder      = 1,
        LowOrder       = 2,
        MixedOrder     = 3
    };

    // Floating-point mode of the GPU pair kernels. FP64 is the reference arithmetic;
    // Mixed evaluates the pair arithmetic in FP32 and accumulates in FP64.
    enum class Precision : int
    {
        FP64  = 0,
        Mixed = 1
    };

    struct ClusterTreeSettings
    {
        int split_threshold     = 1;
        int thread_count        = 1;
        int parallel_perc_depth = 5;
        // Build the cluster tree and the block cluster tree on the GPU (rp_build_topology) instead of the
        // host; same results. No reference counterpart.
        bool device_build       = false;
    };

    struct BlockClusterTreeSettings
    {
        double far_field_separation_parameter    = 0.5;
        double near_field_separation_parameter   = 10.;
        double near_field_intersection_parameter = 10000000000.;
        int    max_refinement                    = 30;
    };

    // Copied into BlockClusterTreeSettings when the block cluster tree is built
    // (src/SimplicialMesh.hpp:592-595).
    struct AdaptivitySettings
    {
        int    max_refinement     = 30;
        double theta              = 10.;
        double intersection_theta = 10000000000.;
    };

} // namespace Repulsor
