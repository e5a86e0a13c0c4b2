#pragma once
// TangentPointMetric0 with the reference's public surface (src/Metrics/TangentPointMetric0.hpp:29-126 on
// MetricDimRestricted / MetricBase, src/Metrics/MetricDimRestricted.hpp:28-297, src/Metrics/MetricBase.hpp:34-134):
// MetricValues, MultiplyMetric (pointer and tensor forms), MultiplyPreconditioner, Solve (by default the
// reference's left-preconditioned GMRES with P0), MetricNorm, PreconditionerNorm, reachable through
// MetricBase<MeshBase> pointers made by TangentPointMetric0_Factory. The metric blocks are assembled and
// applied on the GPU.

#include "MetricBase.hpp"
#include "SimplicialMesh.hpp"
#include "LearnedPreconditioner.hpp"

#include <cstring>

namespace Repulsor
{
    // Krylov method of TangentPointMetric0::Solve (no reference counterpart: the reference runs GMRES).
    enum class SolverMethod : int
    {
        GMRES = 0,   // left-preconditioned restarted GMRES(100), stopping on |P r|_2 / |P b|_2 (the reference's)
        PCG   = 1    // preconditioned conjugate gradients, stopping on |B - A X|_2 / |B|_2 per column
    };

    template<typename Mesh_T>
    class TangentPointMetric0 : public MetricDimRestricted<Mesh_T, OperatorType::MixedOrder>
    {
    public:
        static constexpr int DOM_DIM = Mesh_T::DOM_DIM;
        static constexpr int AMB_DIM = 3;

        using Base_T           = MetricDimRestricted<Mesh_T, OperatorType::MixedOrder>;
        using MeshBase_T       = typename Base_T::MeshBase_T;
        using ValueContainer_T = typename Base_T::ValueContainer_T;
        using SolverVector_T   = typename Base_T::SolverVector_T;
        using ExtReal          = typename Base_T::ExtReal;
        using Real             = typename Base_T::Real;
        using Int              = typename Base_T::Int;

        using Base_T::MetricValues;
        using Base_T::MultiplyMetric;
        using Base_T::MultiplyPreconditioner;
        using Base_T::Solve;

    protected:
        double q, p, s;
        Precision precision = Precision::FP64;
        int pc_precision = -1;   // precision of the pseudo-Laplacian values inside the preconditioner; -1: as precision

    public:
        explicit TangentPointMetric0( const double p_ ) : q( p_ ), p( 2 * p_ ), s( (2 * p_ - DOM_DIM) / p_ ) {}
        TangentPointMetric0( const double q_, const double p_ ) : q( q_ ), p( p_ ), s( (p_ - DOM_DIM) / q_ ) {}

        void SetPrecision( Precision prec ) { precision = prec; }
        Precision GetPrecision() const { return precision; }

        // Storage and pair arithmetic of the pseudo-Laplacian inside the preconditioner (no reference
        // counterpart). PCG and GMRES need only a fixed symmetric positive definite preconditioner, so FP32
        // values there leave the operator, its FP64 residual and the converged solution's accuracy as they
        // are; the metric itself keeps SetPrecision. Unset, it follows SetPrecision, with MixedRefined
        // (FP32 products inside the CG iterations) giving Mixed.
        void SetPreconditionerPrecision( Precision prec )
        {
            pc_precision = prec == Precision::MixedRefined ? (int)Precision::Mixed : (int)prec;
        }
        Precision GetPreconditionerPrecision() const
        {
            if( pc_precision >= 0 ) { return (Precision)pc_precision; }
            return precision == Precision::MixedRefined ? Precision::Mixed : precision;
        }

        std::string ClassName() const override
        {
            char buf[128];
            std::snprintf( buf, sizeof(buf), "TangentPointMetric0<%d,3>(%g,%g,%g)%s", DOM_DIM, q, p, s, PrecisionTag( precision ) );
            return std::string(buf);
        }

        // Krylov method and stopping measure of Solve without explicit solver flags (default: GMRES, as the
        // reference).
        void SetSolverMethod( const SolverMethod m ) { solver_method = m; }
        SolverMethod GetSolverMethod() const { return solver_method; }

        // Assembles the metric blocks on the device (cached until SemiStaticUpdate); the reference's
        // MetricValues(M) assembles and caches them in the mesh (MetricDimRestricted.hpp:28-50).
        bool AssembleMetric( const Mesh_T & M ) const
        {
            try
            {
                rp_ctx * c = M.Device();
                const std::string tag = ClassName();
                auto it = M.device_state.find( "metric" );
                if( it != M.device_state.end() && M.device_state.count( "metric:" + tag ) ) { return true; }
                ScopedTimer timer( "TangentPointMetric0::MetricValues" );
                if( !Check( rp_tp0_compute( c, RP_METRIC, q, p, (int)precision, nullptr ), ClassName() + "::MetricValues" ) ) { return false; }
                for( auto k = M.device_state.begin(); k != M.device_state.end(); )
                {
                    if( k->first.rfind( "metric", 0 ) == 0 ) { k = M.device_state.erase(k); } else { ++k; }
                }
                M.device_state["metric"] = 1;
                M.device_state["metric:" + tag] = 1;
                return true;
            }
            catch( const std::exception & e ) { eprint( ClassName() + "::MetricValues: " + e.what() ); return false; }
        }

        // Metric blocks in the reference's layout: NF values over the full near CSR (2 per nnz),
        // NF diagonal (2 per primitive), FF values over the far CSR, FF diagonal.
        bool DownloadMetricValues( const Mesh_T & M, double * nf_csr, double * nf_diag, double * ff_vals, double * ff_diag ) const
        {
            if( !AssembleMetric( M ) ) { return false; }
            const auto & B = M.GetBlockClusterTree();
            std::vector<double> tile( (size_t)std::max<LInt>( 1, 2 * B.tiles.val_ptr.back() ) );
            if( !Check( rp_download_metric( M.Device(), nf_csr ? tile.data() : nullptr, nf_diag, ff_vals, ff_diag ), ClassName() ) ) { return false; }
            if( nf_csr ) { M.NearTileValuesToCSR( tile.data(), B.NearCSR(), nf_csr ); }
            return true;
        }

        // Host copy of all metric values in the reference's layout (MetricValues through the base class).
        bool DownloadMetricValues( const Mesh_T & M, ValueContainer_T & out ) const override
        {
            try
            {
                const auto & B = M.GetBlockClusterTree();
                const LInt T = M.SimplexCount(), C = M.GetClusterTree().ClusterCount();
                out.NF.assign( (size_t)( 2 * B.NearCSR().NonzeroCount() ), 0.0 );
                out.NF_diag.assign( (size_t)( 2 * T ), 0.0 );
                out.FF.assign( (size_t)( 2 * B.far.NonzeroCount() ), 0.0 );
                out.FF_diag.assign( (size_t)( 2 * C ), 0.0 );
                if( !DownloadMetricValues( M, out.NF.data(), out.NF_diag.data(), out.FF.data(), out.FF_diag.data() ) ) { return false; }
                const LInt nvf = B.VeryNearCSR().NonzeroCount();
                out.VF.assign( (size_t)( 2 * nvf ), 0.0 );
                out.VF_diag.assign( (size_t)( 2 * T ), 0.0 );
                return nvf == 0 || DownloadVeryNearMetricValues( M, out.VF.data(), out.VF_diag.data() );
            }
            catch( const std::exception & e ) { eprint( ClassName() + "::MetricValues: " + e.what() ); return false; }
        }

        // Very-near-field metric values: vf_vals (a0, a1) per entry of the VeryNear pattern (CSR order),
        // vf_diag T x 2 in tree order (TP0_Traversor::VF_Compute, TP0_Traversor.hpp:280-308).
        bool DownloadVeryNearMetricValues( const Mesh_T & M, double * vf_vals, double * vf_diag ) const
        {
            if( !AssembleMetric( M ) ) { return false; }
            return Check( rp_download_vf_metric( M.Device(), vf_vals, vf_diag ), ClassName() );
        }

        // Pseudo-Laplacian very-near-field values: vf_vals per VeryNear entry, vf_diag T (tree order).
        bool DownloadVeryNearPseudoLaplacianValues( const Mesh_T & M, double * vf_vals, double * vf_diag ) const
        {
            if( !PseudoLaplacianValues( M ) ) { return false; }
            return Check( rp_download_pl_vf( M.Device(), vf_vals, vf_diag ), ClassName() );
        }

        // Y = alpha * A X + beta * Y (MetricDimRestricted::MultiplyMetric, MetricDimRestricted.hpp:109-139).
        // nrhs must equal 3; otherwise an error is printed and Y is left untouched.
        bool MultiplyMetric( const Mesh_T & M, const double alpha, const double * X, const Int ldX,
                             const double beta, double * Y, const Int ldY, const Int nrhs,
                             const bool VF_flag = true, const bool NF_flag = true, const bool FF_flag = true,
                             const bool device_ptrs = false ) const override
        {
            if( nrhs != AMB_DIM )
            {
                eprint( ClassName() + "::MultiplyMetric: nrhs = " + std::to_string(nrhs) + " != 3 = AMB_DIM. The current implementation only accepts nrhs = AMB_DIM. Doing nothing." );
                return false;
            }
            if( !AssembleMetric( M ) ) { return false; }
            ScopedTimer timer( "TangentPointMetric0::MultiplyMetric" );
            return Check( rp_multiply_metric( M.Device(), alpha, X, ldX, beta, Y, ldY, nrhs,
                                              VF_flag, NF_flag, FF_flag, device_ptrs ? RP_DEVICE_PTR : 0 ),
                          ClassName() + "::MultiplyMetric" );
        }

        // Y = alpha * A32 X + beta * Y with the FP32 copies of a MixedRefined metric: the product the
        // CG iterations of Solve use (no reference counterpart).
        bool MultiplyMetricInner( const Mesh_T & M, const double alpha, const double * X, const Int ldX,
                                  const double beta, double * Y, const Int ldY, const bool device_ptrs = false ) const
        {
            if( !AssembleMetric( M ) ) { return false; }
            return Check( rp_multiply_metric( M.Device(), alpha, X, ldX, beta, Y, ldY, AMB_DIM,
                                              1, RP_NF_FP32_COPIES, 1, device_ptrs ? RP_DEVICE_PTR : 0 ),
                          ClassName() + "::MultiplyMetricInner" );
        }

        // Vector overload (MetricBase.hpp:46-59): X, Y are V x 3.
        bool MultiplyMetric( const Mesh_T & M, const double alpha, const std::vector<double> & X,
                             const double beta, std::vector<double> & Y ) const
        {
            return MultiplyMetric( M, alpha, X.data(), 3, beta, Y.data(), 3, 3 );
        }

        // Pseudo-Laplacian L_{2-s} of the preconditioner (the reference's member
        // PseudoLaplacian<Mesh,false> pseudo_lap(2 - s), TangentPointMetric0.hpp:42,53), assembled on the
        // device and cached until SemiStaticUpdate.
        bool PseudoLaplacianValues( const Mesh_T & M ) const
        {
            try
            {
                rp_ctx * c = M.Device();
                char key[96];
                const int pl_prec = (int)GetPreconditionerPrecision();
                std::snprintf( key, sizeof(key), "pl:%.17g:%d", 2.0 - s, pl_prec );
                if( M.device_state.count( key ) ) { return true; }
                ScopedTimer timer( "PseudoLaplacian::MetricValues" );
                if( !Check( rp_pl_compute( c, 2.0 - s, pl_prec ), ClassName() + "::PseudoLaplacianValues" ) ) { return false; }
                for( auto k = M.device_state.begin(); k != M.device_state.end(); )
                {
                    if( k->first.rfind( "pl:", 0 ) == 0 ) { k = M.device_state.erase(k); } else { ++k; }
                }
                M.device_state[key] = 1;
                return true;
            }
            catch( const std::exception & e ) { eprint( ClassName() + "::PseudoLaplacianValues: " + e.what() ); return false; }
        }

        // Y = alpha * H^-1 L_{2-s} H^-1 X + beta * Y (TangentPointMetric0::multiply_preconditioner,
        // TangentPointMetric0.hpp:88-126; nrhs check MetricDimRestricted.hpp:173-195).
        bool MultiplyPreconditioner( const Mesh_T & M, const double alpha, const double * X, const Int ldX,
                                     const double beta, double * Y, const Int ldY, const Int nrhs,
                                     const bool device_ptrs = false ) const override
        {
            if( nrhs != AMB_DIM )
            {
                eprint( ClassName() + "::MultiplyPreconditioner: nrhs = " + std::to_string(nrhs) + " != 3 = AMB_DIM. The current implementation only accepts nrhs = AMB_DIM. Doing nothing." );
                return false;
            }
            if( !PreparePreconditioner( M ) ) { return false; }
            ScopedTimer timer( "TangentPointMetric0::MultiplyPreconditioner" );
            return Check( rp_multiply_preconditioner( M.Device(), alpha, X, ldX, beta, Y, ldY, nrhs, device_ptrs ? RP_DEVICE_PTR : 0 ),
                          ClassName() + "::MultiplyPreconditioner" );
        }

        // X = alpha * A^{-1} B + beta * X on the GPU (MetricDimRestricted::Solve, MetricDimRestricted.hpp:232-289).
        // solver_flags (RP_SOLVE_*); -1 takes the object's SolverMethod: GMRES (default, the reference's
        // left-preconditioned GMRES, stopping on |P r|_2 / |P b|_2) or PCG (stopping on |B - A X|_2 / |B|_2 per
        // column). The preconditioner is P0 = H^-1 L_{2-s} H^-1 unless SetPreconditionerDiagonalWeight or a
        // learned model is set. Returns true iff the solve met the tolerance; on max_iter without convergence
        // it prints a warning, keeps the last iterate in X and returns false (Solver_Status: RP_ESTAGNATION).
        // The iteration count and per-column relative residuals are kept for Solver_IterationCount /
        // Solver_RelativeResiduals.
        bool Solve( const Mesh_T & M, const double alpha, const double * B, const Int ldB,
                    const double beta, double * X, const Int ldX, const Int nrhs,
                    const Int max_iter, const double tolerance, const bool device_ptrs = false,
                    const int solver_flags = -1 ) const override
        {
            const int flags = solver_flags >= 0 ? solver_flags : ( solver_method == SolverMethod::GMRES ? RP_SOLVE_GMRES : 0 );
            const bool use_preconditioner = !( flags & RP_SOLVE_NO_PRECONDITIONER );
            status = RP_EINVAL;
            if( nrhs != AMB_DIM )
            {
                eprint( ClassName() + "::Solve: nrhs = " + std::to_string(nrhs) + " != 3 = AMB_DIM. The current implementation only accepts nrhs = AMB_DIM. Doing nothing." );
                return false;
            }
            if( !AssembleMetric( M ) ) { return false; }
            if( use_preconditioner && !PreparePreconditioner( M ) ) { return false; }
            ScopedTimer timer( "TangentPointMetric0::Solve" );
            const bool learned = use_preconditioner && !lp_params.empty();
            const bool can_fall_back = learned && ( beta == 0.0 || !device_ptrs );
            std::vector<double> X0;
            if( can_fall_back && beta != 0.0 ) { X0.assign( X, X + (size_t)ldX * (size_t)( M.VertexCount() - 1 ) + 3 ); }
            int it = 0;
            double rel[3] = { 0, 0, 0 };
            // A learned PCG solve runs under the residual guard: by iteration lp_guard_iter it must have reached
            // the residual that the base preconditioner reaches by lp_guard_iter / LP_GUARD_SLOWDOWN iterations
            // (base_rate per iteration); otherwise it stops early and the base preconditioner takes over.
            const bool guarded = can_fall_back && !( flags & RP_SOLVE_GMRES ) && lp_guard_iter > 0;
            if( guarded && !Check( rp_set_solve_guard( M.Device(), lp_guard_iter, std::pow( base_rate, lp_guard_iter / LP_GUARD_SLOWDOWN ) ),
                                   ClassName() + "::Solve" ) ) { return false; }
            const int st = rp_solve( M.Device(), alpha, B, ldB, beta, X, ldX, nrhs, max_iter, tolerance,
                                     flags, device_ptrs ? RP_DEVICE_PTR : 0, &it, rel );
            const std::string why = st == RP_OK ? std::string() : std::string( rp_last_error() );
            if( guarded ) { rp_set_solve_guard( M.Device(), 0, 0.0 ); }
            SetRecord( it, rel );
            fell_back = false;
            status = st;
            if( st == RP_ESTAGNATION && can_fall_back )
            {
                // the two-level preconditioner the learned scalings wrap is the fallback
                wprint( ClassName() + "::Solve: learned preconditioner too slow (" + why + "); falling back to P0 + w diag(A)^-1" );
                if( !X0.empty() ) { std::copy( X0.begin(), X0.end(), X ); }
                TangentPointMetric0 plain( q, p );
                plain.SetPrecision( precision );
                plain.pc_precision = pc_precision;
                plain.solver_method = solver_method;
                plain.SetPreconditionerDiagonalWeight( -1.0 );
                const bool ok = plain.Solve( M, alpha, B, ldB, beta, X, ldX, nrhs, max_iter, tolerance, device_ptrs, solver_flags );
                iter = it + plain.Solver_IterationCount();
                rel_residuals = plain.Solver_RelativeResiduals();
                status = plain.Solver_Status();
                fell_back = true;
                // the base preconditioner's measured rate becomes the guard's reference
                const int pit = plain.Solver_IterationCount();
                double worst = 0.0;
                for( Int k = 0; k < rel_residuals.Size(); ++k ) { worst = std::max( worst, rel_residuals[k] ); }
                if( ok && pit > 0 && worst > 0.0 ) { base_rate = std::pow( worst, 1.0 / pit ); }
                return ok;
            }
            if( st == RP_ESTAGNATION )
            {
                wprint( ClassName() + "::Solve: " + rp_last_error() );
                return false;
            }
            return Check( st, ClassName() + "::Solve" );
        }

        // Whether the last Solve had to fall back from the learned preconditioner.
        bool Solver_FellBack() const { return fell_back; }

        // Status of the last Solve: RP_OK (converged), RP_ESTAGNATION (max_iter reached; X holds the last
        // iterate) or the error status of a failed call.
        int Solver_Status() const { return status; }

        // Residual guard of learned solves (no reference counterpart): the iteration at which a learned PCG
        // solve is compared with the base preconditioner's rate (0 disables the early fallback; the fallback
        // at max_iter stays), and that reference rate (relative residual reduction per iteration; updated by
        // every fallback solve).
        void SetLearnedPreconditionerGuard( const int check_iteration, const double reference_rate = 0.5 )
        {
            lp_guard_iter = std::max( 0, check_iteration );
            base_rate = reference_rate;
        }
        int LearnedPreconditionerGuardIteration() const { return lp_guard_iter; }
        double LearnedPreconditionerReferenceRate() const { return base_rate; }

    protected:
        using Base_T::iter;
        using Base_T::rel_residuals;
        mutable bool fell_back = false;
        // A learned solve may take LP_GUARD_SLOWDOWN times the iterations of the base preconditioner before
        // the guard hands over.
        static constexpr double LP_GUARD_SLOWDOWN = 1.5;
        int lp_guard_iter = 12;
        mutable double base_rate = 0.5;
        mutable int status = RP_OK;
        SolverMethod solver_method = SolverMethod::GMRES;

        void SetRecord( const int it, const double rel[3] ) const
        {
            iter = it;
            rel_residuals = SolverVector_T( 3 );
            for( int k = 0; k < 3; ++k ) { rel_residuals[k] = rel[k]; }
        }

        double pc_diagonal_weight = 0.0;
        std::vector<float> lp_params;      // learned preconditioner model (empty: none)
        int lp_nfeat = 0;
        bool lp_bf16 = true;

        std::string LearnedKey() const
        {
            uint64_t h = 1469598103934665603ull;
            for( float x : lp_params ) { uint32_t b; std::memcpy( &b, &x, 4 ); h = ( h ^ b ) * 1099511628211ull; }
            return "lp:" + std::to_string( h ) + ( lp_bf16 ? ":bf16" : ":fp32" );
        }

        bool PreparePreconditioner( const Mesh_T & M ) const
        {
            try { M.DeviceH1(); }
            catch( const std::exception & e ) { eprint( ClassName() + "::MultiplyPreconditioner: " + e.what() ); return false; }
            if( !PseudoLaplacianValues( M ) ) { return false; }
            const bool learned = !lp_params.empty();
            if( ( learned || pc_diagonal_weight != 0.0 ) && !AssembleMetric( M ) ) { return false; }
            rp_ctx * c = M.Device();
            // the device holds at most one learned model: switch it on or off as this object needs
            std::string have;
            for( auto & kv : M.device_state ) { if( kv.first.rfind( "lp:", 0 ) == 0 ) { have = kv.first; } }
            const std::string want = learned ? LearnedKey() : std::string();
            if( have != want )
            {
                if( !have.empty() ) { M.device_state.erase( have ); }
                if( learned )
                {
                    const std::vector<float> feats = LearnedPreconditionerFeatures<DOM_DIM>( M.GetClusterTree() );
                    if( !Check( rp_pc_set_model( c, lp_nfeat, feats.data(), lp_params.data(), lp_bf16 ? 1 : 0 ), ClassName() + "::MultiplyPreconditioner" ) ) { return false; }
                    M.device_state[want] = 1;
                }
                else if( !Check( rp_pc_set_model( c, 0, nullptr, nullptr, 0 ), ClassName() + "::MultiplyPreconditioner" ) ) { return false; }
            }
            return Check( rp_set_preconditioner( c, pc_diagonal_weight ), ClassName() + "::MultiplyPreconditioner" );
        }

    public:
        // Learned preconditioner D^{1/2} (P0 + w diag(A)^-1) D^{1/2} (no counterpart in the reference):
        // params = W1 (64 x nfeat), b1, w2, b2, log w as written by repulsor_amd/train_learned_pc.py.
        void SetLearnedPreconditioner( const std::vector<float> & params, const int nfeat = LEARNED_PC_FEATURES, const bool bf16 = true )
        {
            if( (int)params.size() != LearnedPreconditionerParameterCount( nfeat ) )
            {
                eprint( ClassName() + "::SetLearnedPreconditioner: wrong parameter count" );
                return;
            }
            lp_params = params; lp_nfeat = nfeat; lp_bf16 = bf16;
        }

        bool LoadLearnedPreconditioner( const std::string & path, const bool bf16 = true )
        {
            std::vector<float> p;
            int nf = 0;
            if( !ReadLearnedPreconditioner( path, p, nf ) ) { eprint( ClassName() + "::LoadLearnedPreconditioner: cannot read " + path ); return false; }
            SetLearnedPreconditioner( p, nf, bf16 );
            return true;
        }

        void ClearLearnedPreconditioner() { lp_params.clear(); lp_nfeat = 0; }
        bool HasLearnedPreconditioner() const { return !lp_params.empty(); }

        // One training step on the pairs recorded by the last Solve with RP_SOLVE_RECORD (see
        // rp_pc_learn_step); params / m / v are updated in place.
        bool PreconditionerLearnStep( const Mesh_T & M, const float lr, const int step, float * params, float * m, float * v,
                                      double * loss, double * grad ) const
        {
            if( lp_params.empty() ) { eprint( ClassName() + "::PreconditionerLearnStep: no model set" ); return false; }
            if( !PreparePreconditioner( M ) ) { return false; }
            const bool ok = Check( rp_pc_learn_step( M.Device(), lr, step, params, m, v, loss, grad ), ClassName() + "::PreconditionerLearnStep" );
            // the device model now differs from this object's parameters: force a reload next time
            for( auto k = M.device_state.begin(); k != M.device_state.end(); )
            {
                if( k->first.rfind( "lp:", 0 ) == 0 ) { k = M.device_state.erase(k); } else { ++k; }
            }
            return ok;
        }

        // Vertex scalings sqrt(d_v) of the learned model on M.
        bool LearnedScaling( const Mesh_T & M, double * sqrt_d ) const
        {
            if( lp_params.empty() || !PreparePreconditioner( M ) ) { return false; }
            return Check( rp_pc_download_scaling( M.Device(), sqrt_d ), ClassName() + "::LearnedScaling" );
        }

    public:
        // Preconditioner P = P0 + w diag(A)^-1 (w = 0: the reference's P0; w < 0: w estimated per state
        // as lambda_max(P0 A) / 4). The diagonal term has no counterpart in the reference: P0 alone
        // under-weights the roughest modes, so its condition number grows under mesh refinement.
        void SetPreconditionerDiagonalWeight( const double w ) { pc_diagonal_weight = w; }
        double PreconditionerDiagonalWeight() const { return pc_diagonal_weight; }

        // Weight in use on the mesh's device context (after estimation).
        double PreconditionerWeightInUse( const Mesh_T & M ) const
        {
            double w = 0;
            Check( rp_preconditioner_weight( M.Device(), &w ), ClassName() + "::PreconditionerWeightInUse" );
            return w;
        }

        // Exact diagonal of the assembled metric (V values).
        bool MetricDiagonal( const Mesh_T & M, double * diag ) const
        {
            if( !AssembleMetric( M ) ) { return false; }
            return Check( rp_download_metric_diagonal( M.Device(), diag ), ClassName() + "::MetricDiagonal" );
        }

    };

    // TangentPointMetric0_Factory<MeshBase_T, MIN_DOM_DIM, MAX_DOM_DIM, MIN_AMB_DIM, MAX_AMB_DIM>::Make(dom_dim,
    // amb_dim, q, p) (reference src/Energies/TangentPointEnergy/TP_Factory.hpp:1-80, instantiated for
    // TangentPointMetric0 at src/Metrics/TangentPointMetric0.hpp:156): a MetricBase pointer to
    // TangentPointMetric0<SimplicialMesh<dom_dim,amb_dim,...>>, or nullptr with an error message for
    // dimensions outside the factory's range or not built here (curves and surfaces in R^3).
    template<typename MeshBase_T, int ...N> class TangentPointMetric0_Factory {};

    template<typename Real, typename Int, typename LInt, typename SReal, typename ExtReal,
             int MIN_DOM_DIM_, int MAX_DOM_DIM_, int MIN_AMB_DIM_, int MAX_AMB_DIM_>
    class TangentPointMetric0_Factory<SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>, MIN_DOM_DIM_, MAX_DOM_DIM_, MIN_AMB_DIM_, MAX_AMB_DIM_>
    {
    public:
        using MeshBase_T = SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>;
        using Root_T     = MetricBase<MeshBase_T>;

        static constexpr int MIN_AMB_DIM = std::max( 1, MIN_AMB_DIM_ );
        static constexpr int MAX_AMB_DIM = std::max( 1, MAX_AMB_DIM_ );
        static constexpr int MIN_DOM_DIM = std::max( 0, MIN_DOM_DIM_ );
        static constexpr int MAX_DOM_DIM = std::min( MAX_AMB_DIM - 1, MAX_DOM_DIM_ );

        [[nodiscard]] std::unique_ptr<Root_T> Make( const Int dom_dim, const Int amb_dim, const Real q, const Real p ) const
        {
            if( dom_dim < MIN_DOM_DIM || dom_dim > MAX_DOM_DIM ) { eprint( ClassName() + "::Make: dom_dim is out of range." ); return Error( dom_dim, amb_dim ); }
            if( amb_dim < MIN_AMB_DIM || amb_dim > MAX_AMB_DIM ) { eprint( ClassName() + "::Make: amb_dim is out of range." ); return Error( dom_dim, amb_dim ); }
            if( dom_dim >= amb_dim || amb_dim != 3 ) { return Error( dom_dim, amb_dim ); }
            if constexpr ( MIN_DOM_DIM <= 2 && 2 <= MAX_DOM_DIM )
            {
                if( dom_dim == 2 ) { return std::unique_ptr<Root_T>( new TangentPointMetric0<SimplicialMesh<2,3,Real,Int,LInt,SReal,ExtReal>>( q, p ) ); }
            }
            if constexpr ( MIN_DOM_DIM <= 1 && 1 <= MAX_DOM_DIM )
            {
                if( dom_dim == 1 ) { return std::unique_ptr<Root_T>( new TangentPointMetric0<SimplicialMesh<1,3,Real,Int,LInt,SReal,ExtReal>>( q, p ) ); }
            }
            return Error( dom_dim, amb_dim );
        }

        std::string ClassName() const
        {
            return "TangentPointMetric0_Factory<" + std::to_string( MIN_DOM_DIM ) + "," + std::to_string( MAX_DOM_DIM ) + ","
                 + std::to_string( MIN_AMB_DIM ) + "," + std::to_string( MAX_AMB_DIM ) + ">";
        }

    private:
        std::unique_ptr<Root_T> Error( const Int dom_dim, const Int amb_dim ) const
        {
            eprint( ClassName() + " cannot create TangentPointMetric0 with domain dimension " + std::to_string( dom_dim )
                  + " and ambient dimension " + std::to_string( amb_dim ) + "." );
            return nullptr;
        }
    };

} // namespace Repulsor
