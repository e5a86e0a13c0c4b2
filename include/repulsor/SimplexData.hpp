#pragma once
// Per-simplex data of the tangent-point path: area, hull, barycenter, normal projector and the
// surface-gradient operator. Restates SimplexDataKernel (src/SimplexDataKernel.hpp:92-238) for
// simplices of dimension n in {1,2} in R^3.
//
//   near row (16 for triangles, 13 for edges): [a, hull (n+1)x3 in input vertex order, P packed (6)]
//   far  row (10):                              [a, barycenter (3), P packed (6)]
//   a = charge * (1/n!) * prod diag(chol(df^T df)),  df built from the *sorted* simplex,
//   P = I - df (df^T df)^{-1} df^T,  packed 00,01,02,11,12,22.
//   DiffOp (3 x (n+1), row-major, columns = sorted simplex):  D[k][0] = -sum_l dfdag[l][k],
//   D[k][l+1] = dfdag[l][k],  dfdag = (df^T df)^{-1} df^T.

#include "Common.hpp"

namespace Repulsor
{
    template<int DOM_DIM>
    struct SimplexData
    {
        static constexpr int AMB_DIM   = 3;
        static constexpr int SIZE      = DOM_DIM + 1;
        static constexpr int PROJ_DIM  = 6;
        static constexpr int NEAR_DIM  = 1 + SIZE * AMB_DIM + PROJ_DIM;
        static constexpr int FAR_DIM   = 1 + AMB_DIM + PROJ_DIM;
        static constexpr int HULL_SIZE = SIZE * AMB_DIM;

        Int    simplex   [SIZE];
        Int    s_simplex [SIZE];
        double hull      [SIZE][AMB_DIM];
        double center    [AMB_DIM];
        double P         [AMB_DIM][AMB_DIM];
        double dfdagger  [DOM_DIM][AMB_DIM];
        double a      = 0;
        double charge = 0;

        static double StandardSimplexVolume()
        {
            double f = 1; for( int k = 2; k <= DOM_DIM; ++k ) { f *= k; } return 1.0 / f;
        }

        void Read( cptr<double> V, cptr<Int> S, cptr<double> V_charges, const LInt i )
        {
            for( int l = 0; l < SIZE; ++l ) { simplex[l] = S[(LInt)SIZE * i + l]; s_simplex[l] = simplex[l]; }
            std::sort( s_simplex, s_simplex + SIZE );

            for( int l = 0; l < SIZE; ++l )
            {
                for( int k = 0; k < AMB_DIM; ++k ) { hull[l][k] = V[(LInt)AMB_DIM * simplex[l] + k]; }
            }
            for( int k = 0; k < AMB_DIM; ++k ) { center[k] = hull[0][k]; }
            charge = V_charges ? V_charges[simplex[0]] : 1.0;
            for( int l = 1; l < SIZE; ++l )
            {
                for( int k = 0; k < AMB_DIM; ++k ) { center[k] += hull[l][k]; }
                charge += V_charges ? V_charges[simplex[l]] : 1.0;
            }
            const double nth = 1.0 / SIZE;
            for( int k = 0; k < AMB_DIM; ++k ) { center[k] *= nth; }
            charge *= nth;

            a = charge * StandardSimplexVolume();

            double df[AMB_DIM][DOM_DIM];
            for( int l = 0; l < DOM_DIM; ++l )
            {
                for( int k = 0; k < AMB_DIM; ++k )
                {
                    df[k][l] = V[(LInt)AMB_DIM * s_simplex[l+1] + k] - V[(LInt)AMB_DIM * s_simplex[0] + k];
                }
            }
            // g = df^T df, Cholesky g = U^T U (upper U).
            double g[DOM_DIM][DOM_DIM];
            for( int r = 0; r < DOM_DIM; ++r )
            {
                for( int c = 0; c < DOM_DIM; ++c )
                {
                    double s = 0;
                    for( int k = 0; k < AMB_DIM; ++k ) { s += df[k][r] * df[k][c]; }
                    g[r][c] = s;
                }
            }
            double U[DOM_DIM][DOM_DIM] = {};
            for( int r = 0; r < DOM_DIM; ++r )
            {
                double d = g[r][r];
                for( int k = 0; k < r; ++k ) { d -= U[k][r] * U[k][r]; }
                U[r][r] = std::sqrt(d);
                for( int c = r + 1; c < DOM_DIM; ++c )
                {
                    double s = g[r][c];
                    for( int k = 0; k < r; ++k ) { s -= U[k][r] * U[k][c]; }
                    U[r][c] = s / U[r][r];
                }
            }
            for( int l = 0; l < DOM_DIM; ++l ) { a *= U[l][l]; }

            // dfdagger = g^{-1} df^T by two triangular solves per column.
            for( int k = 0; k < AMB_DIM; ++k )
            {
                double y[DOM_DIM];
                for( int r = 0; r < DOM_DIM; ++r )
                {
                    double s = df[k][r];
                    for( int q = 0; q < r; ++q ) { s -= U[q][r] * y[q]; }
                    y[r] = s / U[r][r];
                }
                for( int r = DOM_DIM; r-- > 0; )
                {
                    double s = y[r];
                    for( int q = r + 1; q < DOM_DIM; ++q ) { s -= U[r][q] * dfdagger[q][k]; }
                    dfdagger[r][k] = s / U[r][r];
                }
            }
            for( int r = 0; r < AMB_DIM; ++r )
            {
                for( int c = 0; c < AMB_DIM; ++c )
                {
                    double s = 0;
                    for( int l = 0; l < DOM_DIM; ++l ) { s += df[r][l] * dfdagger[l][c]; }
                    P[r][c] = (r == c ? 1.0 : 0.0) - s;
                }
            }
        }

        void WriteNear( mptr<double> near ) const
        {
            near[0] = a;
            for( int l = 0; l < SIZE; ++l ) { for( int k = 0; k < AMB_DIM; ++k ) { near[1 + AMB_DIM * l + k] = hull[l][k]; } }
            for( int k = 0; k < PROJ_DIM; ++k ) { near[1 + HULL_SIZE + k] = P[ tri_i<3>(k) ][ tri_j<3>(k) ]; }
        }

        void WriteFar( mptr<double> far ) const
        {
            far[0] = a;
            for( int k = 0; k < AMB_DIM; ++k ) { far[1 + k] = center[k]; }
            for( int k = 0; k < PROJ_DIM; ++k ) { far[1 + AMB_DIM + k] = P[ tri_i<3>(k) ][ tri_j<3>(k) ]; }
        }

        // 3 x SIZE row-major; columns follow the sorted simplex.
        void WriteDiffOp( mptr<double> D ) const
        {
            for( int k = 0; k < AMB_DIM; ++k )
            {
                D[SIZE * k] = 0;
                for( int l = 0; l < DOM_DIM; ++l )
                {
                    D[SIZE * k    ] -= dfdagger[l][k];
                    D[SIZE * k + l + 1] = dfdagger[l][k];
                }
            }
        }
    };

    // Near/far rows for all simplices (original order), as ComputeNearFarData does
    // (src/SimplicialMesh/ComputeNearFarData.hpp:3-27).
    template<int DOM_DIM>
    inline void ComputeNearFarData(
        cptr<double> V, cptr<Int> S, cptr<double> V_charges, const LInt T,
        mptr<double> P_near, mptr<double> P_far, const int threads )
    {
        using K = SimplexData<DOM_DIM>;
        ParallelFor( T, threads, [&]( LInt i )
        {
            K ker;
            ker.Read( V, S, V_charges, i );
            ker.WriteNear( P_near + (LInt)K::NEAR_DIM * i );
            ker.WriteFar ( P_far  + (LInt)K::FAR_DIM  * i );
        });
    }

} // namespace Repulsor
