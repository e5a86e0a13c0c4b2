#pragma once
// TangentPointEnergy0 with the reference's public surface (src/Energies/TangentPointEnergy0.hpp:35-86 on
// EnergyDimRestricted / EnergyBase, src/Energies/EnergyBase.hpp:40-121): Value(M), Differential(M) (cached),
// Differential(M, alpha, beta, Y, ldY), Differential(M, Y, addto), reachable through EnergyBase<MeshBase>
// pointers made by TangentPointEnergy0_Factory (src/Energies/TangentPointEnergy/TP_Factory.hpp:1-80).
// The near/far kernels, derivative collection and chain rule run on the GPU through librepulsor_hip.so.

#include "EnergyBase.hpp"
#include "SimplicialMesh.hpp"

namespace Repulsor
{
    template<typename Mesh_T>
    class TangentPointEnergy0 : public EnergyDimRestricted<Mesh_T>
    {
    public:
        static constexpr int DOM_DIM = Mesh_T::DOM_DIM;

        using Base_T            = EnergyDimRestricted<Mesh_T>;
        using MeshBase_T        = typename Base_T::MeshBase_T;
        using CotangentVector_T = typename MeshBase_T::CotangentVector_T;

        using Base_T::Value;
        using Base_T::Differential;

    protected:
        double q, p;
        Precision precision = Precision::FP64;
        using Base_T::last_ok;

    public:
        // (p): q = p, p = 2p  (src/Energies/TangentPointEnergy0.hpp:35-39)
        explicit TangentPointEnergy0( const double p_ ) : q( p_ ), p( 2 * p_ ) {}
        TangentPointEnergy0( const double q_, const double p_ ) : q( q_ ), p( p_ ) {}

        void SetPrecision( Precision prec ) { precision = prec; }
        Precision GetPrecision() const { return precision; }
        double Q() const { return q; }
        double P() const { return p; }

        std::string ClassName() const override
        {
            char buf[128];
            std::snprintf( buf, sizeof(buf), "TangentPointEnergy0<%d,3>(%g,%g)%s", DOM_DIM, q, p, PrecisionTag( precision ) );
            return std::string(buf);
        }

        // Energy and differential into a device buffer (no host round trip; no reference counterpart).
        double DifferentialDevice( const Mesh_T & M, const double alpha, const double beta, double * Y_dev, const LInt ldY ) const
        {
            double en = 0;
            if( !compute( M, RP_ENERGY | RP_DIFF, &en ) ) { return 0; }
            if( !Check( rp_assemble_vertex_gradient( M.Device(), alpha, beta, Y_dev, ldY, RP_DEVICE_PTR ), ClassName() + "::DifferentialDevice" ) ) { return 0; }
            return en;
        }

        // Moment derivatives of the last differential (ClusterTree::PrimitiveDNearFieldData/DFarFieldData).
        bool MomentDerivatives( const Mesh_T & M, double * P_D_near, double * P_D_far ) const
        {
            return Check( rp_download_derivatives( M.Device(), P_D_near, P_D_far ), ClassName() + "::MomentDerivatives" );
        }

    protected:
        double value( cref<Mesh_T> M ) const override
        {
            double en = 0;
            if( !compute( M, RP_ENERGY, &en ) ) { last_ok = false; return 0; }
            return en;
        }

        // Energy and moment derivatives on the device; the mesh's Assemble_ClusterTree_Derivatives then
        // applies the chain rule and the vertex assembly.
        double differential( cref<Mesh_T> M ) const override
        {
            double en = 0;
            if( !compute( M, RP_ENERGY | RP_DIFF, &en ) ) { last_ok = false; return 0; }
            return en;
        }

        bool compute( const Mesh_T & M, unsigned flags, double * en ) const
        {
            try
            {
                rp_ctx * c = M.Device();
                return Check( rp_tp0_compute( c, flags, q, p, (int)precision, en ), ClassName() );
            }
            catch( const std::exception & e ) { eprint( ClassName() + ": " + e.what() ); return false; }
        }
    };

    // TangentPointEnergy0_Factory<MeshBase_T, MIN_DOM_DIM, MAX_DOM_DIM, MIN_AMB_DIM, MAX_AMB_DIM>::Make(dom_dim,
    // amb_dim, q, p) (reference src/Energies/TangentPointEnergy/TP_Factory.hpp:1-80, instantiated for
    // TangentPointEnergy0 at src/Energies/TangentPointEnergy0.hpp:104): an EnergyBase pointer to
    // TangentPointEnergy0<SimplicialMesh<dom_dim,amb_dim,...>>, or nullptr with an error message for
    // dimensions outside the factory's range or not built here (curves and surfaces in R^3).
    template<typename MeshBase_T, int ...N> class TangentPointEnergy0_Factory {};

    template<typename Real, typename Int, typename LInt, typename SReal, typename ExtReal,
             int MIN_DOM_DIM_, int MAX_DOM_DIM_, int MIN_AMB_DIM_, int MAX_AMB_DIM_>
    class TangentPointEnergy0_Factory<SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>, MIN_DOM_DIM_, MAX_DOM_DIM_, MIN_AMB_DIM_, MAX_AMB_DIM_>
    {
    public:
        using MeshBase_T = SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>;
        using Root_T     = EnergyBase<MeshBase_T>;

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
                if( dom_dim == 2 ) { return std::unique_ptr<Root_T>( new TangentPointEnergy0<SimplicialMesh<2,3,Real,Int,LInt,SReal,ExtReal>>( q, p ) ); }
            }
            if constexpr ( MIN_DOM_DIM <= 1 && 1 <= MAX_DOM_DIM )
            {
                if( dom_dim == 1 ) { return std::unique_ptr<Root_T>( new TangentPointEnergy0<SimplicialMesh<1,3,Real,Int,LInt,SReal,ExtReal>>( q, p ) ); }
            }
            return Error( dom_dim, amb_dim );
        }

        std::string ClassName() const
        {
            return "TangentPointEnergy0_Factory<" + std::to_string( MIN_DOM_DIM ) + "," + std::to_string( MAX_DOM_DIM ) + ","
                 + std::to_string( MIN_AMB_DIM ) + "," + std::to_string( MAX_AMB_DIM ) + ">";
        }

    private:
        std::unique_ptr<Root_T> Error( const Int dom_dim, const Int amb_dim ) const
        {
            eprint( ClassName() + " cannot create TangentPointEnergy0 with domain dimension " + std::to_string( dom_dim )
                  + " and ambient dimension " + std::to_string( amb_dim ) + "." );
            return nullptr;
        }
    };

} // namespace Repulsor
