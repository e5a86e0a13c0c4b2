// A user program written against include/repulsor the way the reference's drivers are written against
// the reference (call sequence of Example2/main.cpp:37-147): polymorphic SimplicialMeshBase /
// EnergyBase / MetricBase types, the three factories, Make_FromFile on a mesh in the reference's text
// format, Differential(M, ptr), Solve(..., max_iter = 100, tolerance = 1e-5) with the default solver
// (left-preconditioned GMRES with P0), MaximumSafeStepSize, SemiStaticUpdate and Value.
//
// Usage: reference_call_sequence <mesh file> <output file>
// Output (binary, little-endian): int32 V, int32 iterations, 3 x double relative residuals, double energy,
// double step size t, double energy after the step, V x 3 doubles differential, V x 3 doubles gradient.

#include "repulsor/Repulsor.hpp"

#include <cstdio>
#include <iostream>

using Int  = int;
using LInt = std::size_t;
using Real = double;

int main( int argc, char ** argv )
{
    if( argc < 3 ) { std::fprintf( stderr, "usage: %s <mesh file> <output file>\n", argv[0] ); return 2; }

    using Mesh_T   = Repulsor::SimplicialMeshBase<Real,Int,LInt>;
    using Energy_T = Repulsor::EnergyBase<Mesh_T>;
    using Metric_T = Repulsor::MetricBase<Mesh_T>;

    const Int thread_count = 8;
    const Real q = 6;
    const Real p = 12;

    Repulsor::SimplicialMesh_Factory<Mesh_T,2,2,3,3> mesh_factory;
    Repulsor::TangentPointEnergy0_Factory<Mesh_T,2,2,3,3> TPE_factory;
    Repulsor::TangentPointMetric0_Factory<Mesh_T,2,2,3,3> TPM_factory;

    std::unique_ptr<Energy_T> tpe_ptr = TPE_factory.Make( 2, 3, q, p );
    std::unique_ptr<Metric_T> tpm_ptr = TPM_factory.Make( 2, 3, q, p );
    if( !tpe_ptr || !tpm_ptr ) { return 1; }

    const auto & tpe = *tpe_ptr;
    const auto & tpm = *tpm_ptr;

    std::unique_ptr<Mesh_T> M_ptr = mesh_factory.Make_FromFile( argv[1], thread_count );
    if( !M_ptr ) { return 1; }
    auto & M = *M_ptr;

    M.cluster_tree_settings.split_threshold                      = 2;
    M.cluster_tree_settings.thread_count                         = 0;
    M.block_cluster_tree_settings.far_field_separation_parameter = 0.25;
    M.adaptivity_settings.theta                                  = 10.0;

    Real en;
    Mesh_T::CotangentVector_T diff( M.VertexCount(), M.AmbDim() );
    en = tpe.Differential( M, diff.data() );
    std::cout << "energy " << en << std::endl;

    const Int  max_iter           = 100;
    const Real relative_tolerance = 0.00001;

    Mesh_T::TangentVector_T gradient( M.VertexCount(), M.AmbDim() );
    tpm.Solve( M,
        Real(1), diff.data(),     M.AmbDim(),
        Real(0), gradient.data(), M.AmbDim(),
        M.AmbDim(),
        max_iter, relative_tolerance
    );
    std::cout << "iterations " << tpm.Solver_IterationCount() << ", relative residuals " << tpm.Solver_RelativeResiduals() << std::endl;

    Mesh_T::TangentVector_T downward_gradient = gradient;
    downward_gradient *= -1.;

    const Real t = M.MaximumSafeStepSize( downward_gradient.data(), 1. );
    std::cout << "collision-free step size " << t << std::endl;

    Mesh_T::TangentVector_T X = downward_gradient;
    X *= t;
    X += M.VertexCoordinates();
    M.SemiStaticUpdate( X.data() );

    Real en_new = tpe.Value( M );
    std::cout << "energy after the step " << en_new << " (decreased: " << ( en_new < en ) << ")" << std::endl;

    std::FILE * f = std::fopen( argv[2], "wb" );
    if( !f ) { std::fprintf( stderr, "cannot write %s\n", argv[2] ); return 1; }
    const int32_t V = M.VertexCount();
    const int32_t it = tpm.Solver_IterationCount();
    const auto rr = tpm.Solver_RelativeResiduals();
    double rel[3] = { 0, 0, 0 };
    for( Int k = 0; k < rr.Size() && k < 3; ++k ) { rel[k] = rr[k]; }
    bool ok = std::fwrite( &V, sizeof(V), 1, f ) == 1 && std::fwrite( &it, sizeof(it), 1, f ) == 1
           && std::fwrite( rel, sizeof(double), 3, f ) == 3 && std::fwrite( &en, sizeof(double), 1, f ) == 1
           && std::fwrite( &t, sizeof(double), 1, f ) == 1 && std::fwrite( &en_new, sizeof(double), 1, f ) == 1
           && std::fwrite( diff.data(), sizeof(double), diff.size(), f ) == diff.size()
           && std::fwrite( gradient.data(), sizeof(double), gradient.size(), f ) == gradient.size();
    ok = ( std::fclose( f ) == 0 ) && ok;
    return ok ? 0 : 1;
}
