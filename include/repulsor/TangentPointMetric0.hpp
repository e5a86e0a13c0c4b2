#pragma once
// TangentPointMetric0 with the reference's public surface
// (src/Metrics/TangentPointMetric0.hpp:29-126, src/Metrics/MetricBase.hpp:34-134,
//  src/Metrics/MetricDimRestricted.hpp:28-297): MetricValues, MultiplyMetric (pointer and
// vector overloads), MetricNorm. The metric blocks are assembled and applied on the GPU.

#include "SimplicialMesh.hpp"
#include "LearnedPreconditioner.hpp"

#include <cstring>

namespace Repulsor
{
    template<typename Mesh_T>
    class TangentPointMetric0
    {
    public:
        static constexpr int DOM_DIM = Mesh_T::DOM_DIM;
        static constexpr int AMB_DIM = 3;

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

        std::string ClassName() const
        {
            char buf[128];
            std::snprintf( buf, sizeof(buf), "TangentPointMetric0<%d,3>(%g,%g,%g)%s", DOM_DIM, q, p, s, PrecisionTag( precision ) );
            return std::string(buf);
        }

        // Assembles the metric blocks on the device (cached until SemiStaticUpdate).
        bool MetricValues( const Mesh_T & M ) const
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
            if( !MetricValues( M ) ) { return false; }
            const auto & B = M.GetBlockClusterTree();
            std::vector<double> tile( (size_t)std::max<LInt>( 1, 2 * B.tiles.val_ptr.back() ) );
            if( !Check( rp_download_metric( M.Device(), nf_csr ? tile.data() : nullptr, nf_diag, ff_vals, ff_diag ), ClassName() ) ) { return false; }
            if( nf_csr ) { M.NearTileValuesToCSR( tile.data(), B.NearCSR(), nf_csr ); }
            return true;
        }

        // Very-near-field metric values: vf_vals (a0, a1) per entry of the VeryNear pattern (CSR order),
        // vf_diag T x 2 in tree order (TP0_Traversor::VF_Compute, TP0_Traversor.hpp:280-308).
        bool DownloadVeryNearMetricValues( const Mesh_T & M, double * vf_vals, double * vf_diag ) const
        {
            if( !MetricValues( M ) ) { return false; }
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
        bool MultiplyMetric( const Mesh_T & M, const double alpha, const double * X, const LInt ldX,
                             const double beta, double * Y, const LInt ldY, const int nrhs,
                             const bool VF_flag = true, const bool NF_flag = true, const bool FF_flag = true,
                             const bool device_ptrs = false ) const
        {
            if( nrhs != AMB_DIM )
            {
                eprint( ClassName() + "::MultiplyMetric: nrhs = " + std::to_string(nrhs) + " != 3 = AMB_DIM. The current implementation only accepts nrhs = AMB_DIM. Doing nothing." );
                return false;
            }
            if( !MetricValues( M ) ) { return false; }
            ScopedTimer timer( "TangentPointMetric0::MultiplyMetric" );
            return Check( rp_multiply_metric( M.Device(), alpha, X, ldX, beta, Y, ldY, nrhs,
                                              VF_flag, NF_flag, FF_flag, device_ptrs ? RP_DEVICE_PTR : 0 ),
                          ClassName() + "::MultiplyMetric" );
        }

        // Y = alpha * A32 X + beta * Y with the FP32 copies of a MixedRefined metric: the product the
        // CG iterations of Solve use (no reference counterpart).
        bool MultiplyMetricInner( const Mesh_T & M, const double alpha, const double * X, const LInt ldX,
                                  const double beta, double * Y, const LInt ldY, const bool device_ptrs = false ) const
        {
            if( !MetricValues( M ) ) { return false; }
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

        // sqrt(<X, A X>) (MetricBase.hpp:81-100).
        double MetricNorm( const Mesh_T & M, const double * X, const LInt ldX, const int nrhs ) const
        {
            return QuadraticNorm( M, X, ldX, nrhs, false );
        }

        // sqrt(<X, P X>) (MetricBase.hpp:102-121).
        double PreconditionerNorm( const Mesh_T & M, const double * X, const LInt ldX, const int nrhs ) const
        {
            return QuadraticNorm( M, X, ldX, nrhs, true );
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
        bool MultiplyPreconditioner( const Mesh_T & M, const double alpha, const double * X, const LInt ldX,
                                     const double beta, double * Y, const LInt ldY, const int nrhs,
                                     const bool device_ptrs = false ) const
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

        // X = alpha * A^+ B + beta * X with P0 on the GPU (MetricDimRestricted::Solve,
        // MetricDimRestricted.hpp:232-289). solver_flags (RP_SOLVE_*): default preconditioned conjugate
        // gradients with the l2 residual; RP_SOLVE_PC_NORM measures in the preconditioner norm;
        // RP_SOLVE_GMRES runs the reference's left-preconditioned GMRES. The iteration count and the
        // per-column relative residuals are kept for Solver_IterationCount / Solver_RelativeResiduals.
        bool Solve( const Mesh_T & M, const double alpha, const double * B, const LInt ldB,
                    const double beta, double * X, const LInt ldX, const int nrhs,
                    const int max_iter, const double tolerance, const bool device_ptrs = false,
                    const int solver_flags = 0 ) const
        {
            const bool use_preconditioner = !( solver_flags & RP_SOLVE_NO_PRECONDITIONER );
            if( nrhs != AMB_DIM )
            {
                eprint( ClassName() + "::Solve: nrhs = " + std::to_string(nrhs) + " != 3 = AMB_DIM. The current implementation only accepts nrhs = AMB_DIM. Doing nothing." );
                return false;
            }
            if( !MetricValues( M ) ) { return false; }
            if( use_preconditioner && !PreparePreconditioner( M ) ) { return false; }
            ScopedTimer timer( "TangentPointMetric0::Solve" );
            const bool learned = use_preconditioner && !lp_params.empty();
            const bool can_fall_back = learned && ( beta == 0.0 || !device_ptrs );
            std::vector<double> X0;
            if( can_fall_back && beta != 0.0 ) { X0.assign( X, X + ldX * ( M.VertexCount() - 1 ) + 3 ); }
            int it = 0;
            double rel[3] = { 0, 0, 0 };
            const int st = rp_solve( M.Device(), alpha, B, ldB, beta, X, ldX, nrhs, max_iter, tolerance,
                                     solver_flags, device_ptrs ? RP_DEVICE_PTR : 0, &it, rel );
            iter = it;
            rel_residuals.assign( rel, rel + 3 );
            fell_back = false;
            if( st == RP_ESTAGNATION && can_fall_back )
            {
                // the reference preconditioner (with the estimated diagonal term) is the fallback
                wprint( ClassName() + "::Solve: learned preconditioner stagnated; falling back to P0 + w diag(A)^-1" );
                if( !X0.empty() ) { std::copy( X0.begin(), X0.end(), X ); }
                TangentPointMetric0 plain( q, p );
                plain.SetPrecision( precision );
                plain.pc_precision = pc_precision;
                plain.SetPreconditionerDiagonalWeight( -1.0 );
                const bool ok = plain.Solve( M, alpha, B, ldB, beta, X, ldX, nrhs, max_iter, tolerance, device_ptrs, solver_flags );
                iter = it + plain.Solver_IterationCount();
                rel_residuals = plain.Solver_RelativeResiduals();
                fell_back = true;
                return ok;
            }
            if( st == RP_ESTAGNATION )
            {
                wprint( ClassName() + "::Solve: " + rp_last_error() );
                return true;
            }
            return Check( st, ClassName() + "::Solve" );
        }

        // Whether the last Solve had to fall back from the learned preconditioner.
        bool Solver_FellBack() const { return fell_back; }

        int Solver_IterationCount() const { return iter; }
        std::vector<double> Solver_RelativeResiduals() const { return rel_residuals; }

    protected:
        mutable int iter = 0;
        mutable std::vector<double> rel_residuals;
        mutable bool fell_back = false;

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
            if( ( learned || pc_diagonal_weight != 0.0 ) && !MetricValues( M ) ) { return false; }
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

    protected:

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
            if( !MetricValues( M ) ) { return false; }
            return Check( rp_download_metric_diagonal( M.Device(), diag ), ClassName() + "::MetricDiagonal" );
        }

    protected:

        double QuadraticNorm( const Mesh_T & M, const double * X, const LInt ldX, const int nrhs, const bool pc ) const
        {
            const LInt V = M.VertexCount();
            std::vector<double> Xc( (size_t)(3 * V) ), Yc( (size_t)(3 * V), 0.0 );
            for( LInt v = 0; v < V; ++v ) { for( int k = 0; k < 3; ++k ) { Xc[(size_t)(3 * v + k)] = X[ldX * v + k]; } }
            const bool ok = pc ? MultiplyPreconditioner( M, 1.0, Xc.data(), 3, 0.0, Yc.data(), 3, nrhs )
                               : MultiplyMetric( M, 1.0, Xc.data(), 3, 0.0, Yc.data(), 3, nrhs );
            if( !ok ) { return 0; }
            double s2 = 0;
            for( size_t k = 0; k < Xc.size(); ++k ) { s2 += Xc[k] * Yc[k]; }
            return std::sqrt( s2 );
        }
    };

} // namespace Repulsor
