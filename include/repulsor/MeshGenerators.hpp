#pragma once
// Seeded synthetic meshes: torus, subdivided icosphere with spherical-harmonic radial noise,
// Fourier-perturbed trefoil. RNG = splitmix64 seeding xoshiro256**, normals by Box-Muller.
// Vertex coordinates are row-major V x 3, simplices row-major T x (n+1), as the reference's
// SimplicialMesh pointer constructor expects (src/SimplicialMesh.hpp:143-185).

#include "Common.hpp"
#include <map>

namespace Repulsor
{
    class Xoshiro256ss
    {
        uint64_t s[4];

        static uint64_t splitmix64( uint64_t & x )
        {
            uint64_t z = (x += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        }

        static uint64_t rotl( const uint64_t x, int k ) { return (x << k) | (x >> (64 - k)); }

        bool   have_spare = false;
        double spare      = 0;

    public:
        explicit Xoshiro256ss( uint64_t seed )
        {
            uint64_t x = seed;
            for( int i = 0; i < 4; ++i ) { s[i] = splitmix64(x); }
        }

        uint64_t Next()
        {
            const uint64_t result = rotl( s[1] * 5, 7 ) * 9;
            const uint64_t t = s[1] << 17;
            s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
            s[2] ^= t;
            s[3] = rotl( s[3], 45 );
            return result;
        }

        // Uniform in (0,1): 53 random bits, never exactly 0.
        double Uniform() { return ( (double)(Next() >> 11) + 0.5 ) * 0x1.0p-53; }

        double Normal()
        {
            if( have_spare ) { have_spare = false; return spare; }
            const double u1 = Uniform();
            const double u2 = Uniform();
            const double r  = std::sqrt( -2.0 * std::log(u1) );
            const double th = 2.0 * M_PI * u2;
            spare = r * std::sin(th);
            have_spare = true;
            return r * std::cos(th);
        }
    };

    struct MeshData
    {
        int dom_dim = 2;
        int amb_dim = 3;
        std::vector<double> coords;     // V x 3
        std::vector<Int>    simplices;  // T x (dom_dim+1)

        LInt VertexCount()  const { return (LInt)coords.size() / amb_dim; }
        LInt SimplexCount() const { return (LInt)simplices.size() / (dom_dim + 1); }
    };

    // Torus x(u,v) = ((R + r cos v) cos u, (R + r cos v) sin u, r sin v) on a periodic Nu x Nv grid.
    // Quad (i,j) is split along the diagonal (i,j)-(i+1,j+1); both triangles are oriented outward.
    // Vertices are jittered by sigma * N(0,1) along the unit normal.
    inline MeshData MakeTorus( int Nu, int Nv, double R, double r, double sigma, uint64_t seed )
    {
        MeshData M;
        M.dom_dim = 2;
        M.coords.resize( (size_t)Nu * Nv * 3 );
        Xoshiro256ss rng( seed );
        for( int i = 0; i < Nu; ++i )
        {
            const double u = 2.0 * M_PI * i / Nu;
            for( int j = 0; j < Nv; ++j )
            {
                const double v  = 2.0 * M_PI * j / Nv;
                const double n0 = std::cos(u) * std::cos(v);
                const double n1 = std::sin(u) * std::cos(v);
                const double n2 = std::sin(v);
                const double xi = (sigma != 0.0) ? sigma * rng.Normal() : 0.0;
                double * X = &M.coords[ 3 * ((size_t)i * Nv + j) ];
                X[0] = (R + r * std::cos(v)) * std::cos(u) + xi * n0;
                X[1] = (R + r * std::cos(v)) * std::sin(u) + xi * n1;
                X[2] = r * std::sin(v) + xi * n2;
            }
        }
        M.simplices.reserve( (size_t)Nu * Nv * 6 );
        auto id = [Nu,Nv]( int i, int j ) { return (Int)( ((i % Nu + Nu) % Nu) * Nv + ((j % Nv + Nv) % Nv) ); };
        for( int i = 0; i < Nu; ++i )
        {
            for( int j = 0; j < Nv; ++j )
            {
                const Int a = id(i,j), b = id(i+1,j), c = id(i+1,j+1), d = id(i,j+1);
                M.simplices.insert( M.simplices.end(), { a, b, c } );
                M.simplices.insert( M.simplices.end(), { a, c, d } );
            }
        }
        return M;
    }

    // Real orthonormal spherical harmonics Y_lm, 0 <= l <= lmax, evaluated at unit vector x.
    // Output order: for l = 0..lmax, m = -l..l. Fully normalized associated Legendre functions
    // by the standard stable recurrences.
    inline void RealSphericalHarmonics( const int lmax, const double * x, std::vector<double> & Y, std::vector<double> & Pbuf )
    {
        const double ct = std::max( -1.0, std::min( 1.0, x[2] ) );
        const double st = std::sqrt( std::max( 0.0, 1.0 - ct * ct ) );
        const double phi = std::atan2( x[1], x[0] );
        const int L = lmax + 1;
        Pbuf.assign( (size_t)L * L, 0.0 );
        auto P = [&]( int l, int m ) -> double & { return Pbuf[(size_t)l * L + m]; };
        // Normalized \bar P_l^m with \int_{-1}^{1} \bar P^2 = 1/(2 pi) * ... ; we use the convention
        // Y_l0 = \bar P_l^0, Y_lm = sqrt(2) \bar P_l^m cos(m phi), Y_l,-m = sqrt(2) \bar P_l^m sin(m phi),
        // \bar P_l^m = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m (no Condon-Shortley phase).
        P(0,0) = std::sqrt( 1.0 / (4.0 * M_PI) );
        for( int m = 1; m <= lmax; ++m )
        {
            P(m,m) = std::sqrt( (2.0 * m + 1.0) / (2.0 * m) ) * st * P(m-1,m-1);
        }
        for( int m = 0; m < lmax; ++m )
        {
            P(m+1,m) = std::sqrt( 2.0 * m + 3.0 ) * ct * P(m,m);
        }
        for( int m = 0; m <= lmax; ++m )
        {
            for( int l = m + 2; l <= lmax; ++l )
            {
                const double a = std::sqrt( (4.0 * l * l - 1.0) / ((double)l * l - (double)m * m) );
                const double b = std::sqrt( ( (double)(l-1) * (l-1) - (double)m * m ) / ( 4.0 * (l-1) * (l-1) - 1.0 ) );
                P(l,m) = a * ( ct * P(l-1,m) - b * P(l-2,m) );
            }
        }
        Y.assign( (size_t)L * L, 0.0 );
        size_t k = 0;
        for( int l = 0; l <= lmax; ++l )
        {
            for( int m = -l; m <= l; ++m )
            {
                if( m == 0 )     { Y[k] = P(l,0); }
                else if( m > 0 ) { Y[k] = std::sqrt(2.0) * P(l,m)  * std::cos(  m * phi ); }
                else             { Y[k] = std::sqrt(2.0) * P(l,-m) * std::sin( -m * phi ); }
                ++k;
            }
        }
    }

    // Icosahedron subdivided `level` times (edge midpoints projected to the unit sphere;
    // V = 10*4^level + 2), radially perturbed: r(x) = 1 + A f(x) / max_v |f(x_v)|,
    // f = sum_{l<=lmax} sum_m c_lm Y_lm(x), c_lm ~ N(0,1) drawn in (l,m) order from `seed`.
    inline MeshData MakeIcosphere( int level, int lmax, double amplitude, uint64_t seed, int threads = 1 )
    {
        MeshData M;
        M.dom_dim = 2;
        const double t = (1.0 + std::sqrt(5.0)) / 2.0;
        std::vector<double> X = {
            -1, t, 0,   1, t, 0,   -1,-t, 0,   1,-t, 0,
             0,-1, t,   0, 1, t,    0,-1,-t,   0, 1,-t,
             t, 0,-1,   t, 0, 1,   -t, 0,-1,  -t, 0, 1 };
        for( size_t v = 0; v < 12; ++v )
        {
            double * p = &X[3*v];
            const double n = std::sqrt( p[0]*p[0] + p[1]*p[1] + p[2]*p[2] );
            p[0] /= n; p[1] /= n; p[2] /= n;
        }
        std::vector<Int> F = {
            0,11,5,  0,5,1,  0,1,7,  0,7,10,  0,10,11,
            1,5,9,   5,11,4, 11,10,2, 10,7,6,  7,1,8,
            3,9,4,   3,4,2,  3,2,6,   3,6,8,   3,8,9,
            4,9,5,   2,4,11, 6,2,10,  8,6,7,   9,8,1 };
        for( int lev = 0; lev < level; ++lev )
        {
            std::map<std::pair<Int,Int>,Int> mid;
            auto midpoint = [&]( Int a, Int b ) -> Int
            {
                const auto key = std::make_pair( std::min(a,b), std::max(a,b) );
                auto it = mid.find(key);
                if( it != mid.end() ) { return it->second; }
                const Int id = (Int)( X.size() / 3 );
                double p[3] = { 0.5*(X[3*a]+X[3*b]), 0.5*(X[3*a+1]+X[3*b+1]), 0.5*(X[3*a+2]+X[3*b+2]) };
                const double n = std::sqrt( p[0]*p[0] + p[1]*p[1] + p[2]*p[2] );
                X.push_back(p[0]/n); X.push_back(p[1]/n); X.push_back(p[2]/n);
                mid.emplace( key, id );
                return id;
            };
            std::vector<Int> G;
            G.reserve( F.size() * 4 );
            for( size_t f = 0; f < F.size() / 3; ++f )
            {
                const Int a = F[3*f], b = F[3*f+1], c = F[3*f+2];
                const Int ab = midpoint(a,b), bc = midpoint(b,c), ca = midpoint(c,a);
                G.insert( G.end(), { a, ab, ca,  b, bc, ab,  c, ca, bc,  ab, bc, ca } );
            }
            F.swap(G);
        }
        const LInt V = (LInt)X.size() / 3;
        if( amplitude != 0.0 && lmax >= 0 )
        {
            Xoshiro256ss rng( seed );
            const int L = lmax + 1;
            std::vector<double> c( (size_t)L * L );
            for( auto & ci : c ) { ci = rng.Normal(); }
            std::vector<double> f( (size_t)V, 0.0 );
            ParallelDo( threads, [&]( int th )
            {
                std::vector<double> Y, Pb;
                const LInt b = (V * th) / threads, e = (V * (th+1)) / threads;
                for( LInt v = b; v < e; ++v )
                {
                    RealSphericalHarmonics( lmax, &X[3*v], Y, Pb );
                    double s = 0;
                    for( size_t k = 0; k < c.size(); ++k ) { s += c[k] * Y[k]; }
                    f[(size_t)v] = s;
                }
            });
            double fmax = 0;
            for( double fv : f ) { fmax = std::max( fmax, std::abs(fv) ); }
            for( LInt v = 0; v < V; ++v )
            {
                const double rr = 1.0 + amplitude * f[(size_t)v] / fmax;
                X[3*v] *= rr; X[3*v+1] *= rr; X[3*v+2] *= rr;
            }
        }
        M.coords.swap(X);
        M.simplices.swap(F);
        return M;
    }

    // Closed polygon: trefoil x(t) = (sin t + 2 sin 2t, cos t - 2 cos 2t, -sin 3t) plus
    // s * sum_{k=1..modes} (alpha_k cos kt + beta_k sin kt), alpha_k, beta_k ~ N(0,1)^3 from `seed`,
    // s chosen so that max_i |perturbation(t_i)| = amplitude. Edges (i, i+1 mod N).
    inline MeshData MakeTrefoil( LInt N, int modes, double amplitude, uint64_t seed )
    {
        MeshData M;
        M.dom_dim = 1;
        Xoshiro256ss rng( seed );
        std::vector<double> alpha( (size_t)modes * 3 ), beta( (size_t)modes * 3 );
        for( int k = 0; k < modes; ++k )
        {
            for( int d = 0; d < 3; ++d ) { alpha[3*k+d] = rng.Normal(); }
            for( int d = 0; d < 3; ++d ) { beta [3*k+d] = rng.Normal(); }
        }
        std::vector<double> pert( (size_t)N * 3, 0.0 );
        double pmax = 0;
        for( LInt i = 0; i < N; ++i )
        {
            const double tt = 2.0 * M_PI * (double)i / (double)N;
            double p[3] = {0,0,0};
            for( int k = 1; k <= modes; ++k )
            {
                const double ck = std::cos(k * tt), sk = std::sin(k * tt);
                for( int d = 0; d < 3; ++d ) { p[d] += alpha[3*(k-1)+d] * ck + beta[3*(k-1)+d] * sk; }
            }
            for( int d = 0; d < 3; ++d ) { pert[3*i+d] = p[d]; }
            pmax = std::max( pmax, std::sqrt( p[0]*p[0] + p[1]*p[1] + p[2]*p[2] ) );
        }
        const double s = (pmax > 0 && modes > 0) ? amplitude / pmax : 0.0;
        M.coords.resize( (size_t)N * 3 );
        M.simplices.resize( (size_t)N * 2 );
        for( LInt i = 0; i < N; ++i )
        {
            const double tt = 2.0 * M_PI * (double)i / (double)N;
            M.coords[3*i+0] = std::sin(tt) + 2.0 * std::sin(2.0 * tt) + s * pert[3*i+0];
            M.coords[3*i+1] = std::cos(tt) - 2.0 * std::cos(2.0 * tt) + s * pert[3*i+1];
            M.coords[3*i+2] = -std::sin(3.0 * tt)                     + s * pert[3*i+2];
            M.simplices[2*i+0] = (Int)i;
            M.simplices[2*i+1] = (Int)((i + 1) % N);
        }
        return M;
    }

} // namespace Repulsor
