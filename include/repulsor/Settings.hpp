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
    // MixedRefined (metric only; energy and differential compute as FP64): FP64 metric values plus
    // FP32 copies of the near-field values; the CG iterations of Solve multiply with the copies and
    // judge convergence by the FP64 residual, replacing the residual and continuing on the FP64
    // values where needed. MultiplyMetric uses the FP64 values.
    enum class Precision : int
    {
        FP64         = 0,
        Mixed        = 1,
        MixedRefined = 2,
        PseudoLaplacianBF16 = 3   // preconditioner only: near pseudo-Laplacian values stored in BF16
    };

    inline const char * PrecisionTag( Precision prec )
    {
        return prec == Precision::Mixed ? "[mixed]" : prec == Precision::MixedRefined ? "[mixed-refined]" : "";
    }

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
