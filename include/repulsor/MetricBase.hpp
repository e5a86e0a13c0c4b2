#pragma once
// Polymorphic metric interface of the public API.
//
// MetricBase<SimplicialMeshBase<...>> (reference src/Metrics/MetricBase.hpp:6-151): MetricValues,
// MultiplyMetric (pointer and tensor forms), MultiplyPreconditioner, Solve, MetricNorm,
// PreconditionerNorm, Solver_IterationCount, Solver_RelativeResiduals.
// MetricDimRestricted<Mesh_T, op_type> (reference src/Metrics/MetricDimRestricted.hpp:6-297): downcasts the
// mesh to the concrete SimplicialMesh type and forwards to the concrete metric; a mesh of another type gets
// an error message and the output is left untouched.
//
// The metric values live on the device. MetricValues returns a host copy in the reference's layout (the
// near and far blocks in the CSR order of the block cluster tree's patterns), made on each call.

#include "SimplicialMeshBase.hpp"

namespace Repulsor
{
    // Host copy of assembled metric values (reference MetricValueContainer, src/FMM/FMM_Configurator.hpp:5-20):
    // per block the pair (a0, a1) in the CSR order of the Near and Far patterns (full symmetric), the
    // block diagonals (2 per primitive / cluster), and the very-near pairs of the VeryNear pattern.
    struct MetricValueContainer
    {
        std::vector<double> NF, NF_diag, FF, FF_diag, VF, VF_diag;
    };

    template<typename MeshBase_T> class MetricBase {};

    template<typename Real, typename Int, typename LInt, typename SReal, typename ExtReal>
    class MetricBase<SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>>
    {
    public:
        using MeshBase_T        = SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>;
        using TangentVector_T   = typename MeshBase_T::TangentVector_T;
        using CotangentVector_T = typename MeshBase_T::CotangentVector_T;
        using ValueContainer_T  = MetricValueContainer;
        using SolverVector_T    = Tensor1<Real,Int>;

        MetricBase() = default;
        virtual ~MetricBase() = default;

    protected:
        mutable Int iter = 0;
        mutable SolverVector_T rel_residuals;

    public:
        virtual mref<ValueContainer_T> MetricValues( cref<MeshBase_T> M ) const = 0;

        // Y = alpha * A X + beta * Y.
        virtual void MultiplyMetric( cref<MeshBase_T> M,
                                     const ExtReal alpha, cptr<ExtReal> X, const Int ldX,
                                     const ExtReal beta,  mptr<ExtReal> Y, const Int ldY,
                                     const Int nrhs,
                                     const bool VF_flag = true, const bool NF_flag = true, const bool FF_flag = true ) const = 0;

        // Y = alpha * A X + beta * Y with tensors (reference MetricBase.hpp:46-59).
        void MultiplyMetric( cref<MeshBase_T> M,
                             const ExtReal alpha, cref<TangentVector_T> X,
                             const ExtReal beta,  mref<CotangentVector_T> Y,
                             const bool VF_flag = true, const bool NF_flag = true, const bool FF_flag = true ) const
        {
            this->MultiplyMetric( M, alpha, X.data(), X.Dim(1), beta, Y.data(), Y.Dim(1), X.Dim(1), VF_flag, NF_flag, FF_flag );
        }

        // Y = alpha * P X + beta * Y.
        virtual void MultiplyPreconditioner( cref<MeshBase_T> M,
                                             const ExtReal alpha, cptr<ExtReal> X, const Int ldX,
                                             const ExtReal beta,  mptr<ExtReal> Y, const Int ldY,
                                             const Int nrhs ) const = 0;

        // X = alpha * A^{-1} B + beta * X by a preconditioned Krylov method; the iteration count and the
        // per-column relative residuals are kept for Solver_IterationCount / Solver_RelativeResiduals.
        virtual void Solve( cref<MeshBase_T> M,
                            const ExtReal alpha, cptr<ExtReal> B, const Int ldB,
                            const ExtReal beta,  mptr<ExtReal> X, const Int ldX,
                            const Int nrhs, const Int max_iter, const Real tolerance ) const = 0;

        // sqrt(<X, A X>) (reference MetricBase.hpp:62-81; meant for tests).
        ExtReal MetricNorm( cref<MeshBase_T> M, cptr<ExtReal> X, const Int ldX, const Int nrhs ) const
        {
            return quadratic( M, X, ldX, nrhs, false );
        }

        // sqrt(<X, P X>) (reference MetricBase.hpp:83-104).
        ExtReal PreconditionerNorm( cref<MeshBase_T> M, cptr<ExtReal> X, const Int ldX, const Int nrhs ) const
        {
            return quadratic( M, X, ldX, nrhs, true );
        }

        Int Solver_IterationCount() const { return iter; }
        SolverVector_T Solver_RelativeResiduals() const { return rel_residuals; }

        static std::string className() { return "MetricBase<double,int32,LInt,double,double>"; }
        virtual std::string ClassName() const { return className(); }

    private:
        ExtReal quadratic( cref<MeshBase_T> M, cptr<ExtReal> X, const Int ldX, const Int nrhs, const bool pc ) const
        {
            const Int V = M.VertexCount();
            Tensor2<ExtReal,Int> Xc( X, V, nrhs, ldX ), Yc( V, nrhs, ExtReal(0) );
            if( pc ) { MultiplyPreconditioner( M, ExtReal(1), Xc.data(), nrhs, ExtReal(0), Yc.data(), nrhs, nrhs ); }
            else     { MultiplyMetric( M, ExtReal(1), Xc.data(), nrhs, ExtReal(0), Yc.data(), nrhs, nrhs ); }
            double s2 = 0;
            for( size_t k = 0; k < Xc.size(); ++k ) { s2 += Xc.data()[k] * Yc.data()[k]; }
            return std::sqrt( s2 );
        }
    };

    template<typename Mesh_T, OperatorType op_type>
    class MetricDimRestricted : public MetricBase<typename Mesh_T::Base_T>
    {
    public:
        using MeshBase_T       = typename Mesh_T::Base_T;
        using Base_T           = MetricBase<MeshBase_T>;
        using ValueContainer_T = typename Base_T::ValueContainer_T;
        using ExtReal          = typename MeshBase_T::ExtReal;
        using Real             = typename MeshBase_T::Real;
        using Int              = typename MeshBase_T::Int;

        using Base_T::MultiplyMetric;

        MetricDimRestricted() = default;
        ~MetricDimRestricted() override = default;

        mref<ValueContainer_T> MetricValues( cref<MeshBase_T> M ) const override
        {
            const Mesh_T * Q = downcast( M, "MetricValues" );
            if( Q == nullptr ) { values = ValueContainer_T(); return values; }
            if( !DownloadMetricValues( *Q, values ) ) { values = ValueContainer_T(); }
            return values;
        }

        void MultiplyMetric( cref<MeshBase_T> M,
                             const ExtReal alpha, cptr<ExtReal> X, const Int ldX,
                             const ExtReal beta,  mptr<ExtReal> Y, const Int ldY,
                             const Int nrhs,
                             const bool VF_flag = true, const bool NF_flag = true, const bool FF_flag = true ) const override
        {
            if( const Mesh_T * Q = downcast( M, "MultiplyMetric" ) )
            {
                (void)MultiplyMetric( *Q, alpha, X, ldX, beta, Y, ldY, nrhs, VF_flag, NF_flag, FF_flag, false );
            }
        }

        void MultiplyPreconditioner( cref<MeshBase_T> M,
                                     const ExtReal alpha, cptr<ExtReal> X, const Int ldX,
                                     const ExtReal beta,  mptr<ExtReal> Y, const Int ldY,
                                     const Int nrhs ) const override
        {
            if( const Mesh_T * Q = downcast( M, "MultiplyPreconditioner" ) )
            {
                (void)MultiplyPreconditioner( *Q, alpha, X, ldX, beta, Y, ldY, nrhs, false );
            }
        }

        void Solve( cref<MeshBase_T> M,
                    const ExtReal alpha, cptr<ExtReal> B, const Int ldB,
                    const ExtReal beta,  mptr<ExtReal> X, const Int ldX,
                    const Int nrhs, const Int max_iter, const Real tolerance ) const override
        {
            if( const Mesh_T * Q = downcast( M, "Solve" ) )
            {
                (void)Solve( *Q, alpha, B, ldB, beta, X, ldX, nrhs, max_iter, tolerance, false, -1 );
            }
        }

        // The concrete metric's forms on the concrete mesh type (the same parameter types as above, so a
        // call with a concrete mesh picks these); they return whether the call succeeded (Solve: whether it
        // converged), take device pointers with device_ptrs, and Solve takes explicit solver flags
        // (RP_SOLVE_*; -1: the metric's default solver).
        virtual bool DownloadMetricValues( const Mesh_T & M, ValueContainer_T & out ) const = 0;
        virtual bool MultiplyMetric( const Mesh_T & M, const ExtReal alpha, cptr<ExtReal> X, const Int ldX,
                                     const ExtReal beta, mptr<ExtReal> Y, const Int ldY, const Int nrhs,
                                     const bool VF_flag = true, const bool NF_flag = true, const bool FF_flag = true,
                                     const bool device_ptrs = false ) const = 0;
        virtual bool MultiplyPreconditioner( const Mesh_T & M, const ExtReal alpha, cptr<ExtReal> X, const Int ldX,
                                             const ExtReal beta, mptr<ExtReal> Y, const Int ldY, const Int nrhs,
                                             const bool device_ptrs = false ) const = 0;
        virtual bool Solve( const Mesh_T & M, const ExtReal alpha, cptr<ExtReal> B, const Int ldB,
                            const ExtReal beta, mptr<ExtReal> X, const Int ldX, const Int nrhs,
                            const Int max_iter, const Real tolerance, const bool device_ptrs = false,
                            const int solver_flags = -1 ) const = 0;

        static std::string className()
        {
            return "MetricDimRestricted<" + std::to_string( Mesh_T::DOM_DIM ) + "," + std::to_string( Mesh_T::AMB_DIM ) + ">";
        }
        std::string ClassName() const override { return className(); }

    protected:
        mutable ValueContainer_T values;

        const Mesh_T * downcast( cref<MeshBase_T> M, const char * what ) const
        {
            const Mesh_T * Q = dynamic_cast<const Mesh_T *>( &M );
            if( Q == nullptr ) { eprint( this->ClassName() + "::" + what + ": Input could not be downcast to compatible type. Doing nothing." ); }
            return Q;
        }
    };

} // namespace Repulsor
