#pragma once
// Exact Euclidean distance between two simplices (points, segments, triangles) in R^3.
//
// The reference decides the near-field MAC between two primitives with GJK
// (src/GJK.hpp:618-766, MultipoleAcceptanceCriterion :1162-1172):
//     separated  <=>  max(r_P^2, r_Q^2) < theta^2 * dist(P,Q)^2,
// where GJK approximates dist. Here dist is computed exactly from closest-point queries
// (Ericson, "Real-Time Collision Detection", 5.1.5 point-triangle and 5.1.9 segment-segment),
// with an explicit edge-triangle crossing test so that intersecting triangles report 0.

#include "Common.hpp"

namespace Repulsor
{
    namespace detail
    {
        struct V3 { double x, y, z; };
        inline V3 operator-( V3 a, V3 b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
        inline V3 operator+( V3 a, V3 b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
        inline V3 operator*( double s, V3 a ) { return { s * a.x, s * a.y, s * a.z }; }
        inline double dot( V3 a, V3 b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
        inline V3 cross( V3 a, V3 b ) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
        inline double clamp01( double t ) { return t < 0 ? 0 : (t > 1 ? 1 : t); }

        inline V3 ClosestPtPointTriangle( V3 p, V3 a, V3 b, V3 c )
        {
            const V3 ab = b - a, ac = c - a, ap = p - a;
            const double d1 = dot(ab, ap), d2 = dot(ac, ap);
            if( d1 <= 0 && d2 <= 0 ) { return a; }
            const V3 bp = p - b;
            const double d3 = dot(ab, bp), d4 = dot(ac, bp);
            if( d3 >= 0 && d4 <= d3 ) { return b; }
            const double vc = d1 * d4 - d3 * d2;
            if( vc <= 0 && d1 >= 0 && d3 <= 0 ) { const double v = d1 / (d1 - d3); return a + v * ab; }
            const V3 cp = p - c;
            const double d5 = dot(ab, cp), d6 = dot(ac, cp);
            if( d6 >= 0 && d5 <= d6 ) { return c; }
            const double vb = d5 * d2 - d1 * d6;
            if( vb <= 0 && d2 >= 0 && d6 <= 0 ) { const double w = d2 / (d2 - d6); return a + w * ac; }
            const double va = d3 * d6 - d5 * d4;
            if( va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0 )
            {
                const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
                return b + w * (c - b);
            }
            const double denom = 1.0 / (va + vb + vc);
            const double v = vb * denom, w = vc * denom;
            return a + v * ab + w * ac;
        }

        inline double SegSegDist2( V3 p1, V3 q1, V3 p2, V3 q2 )
        {
            const V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
            const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
            double s, t;
            const double EPS = 1e-300;
            if( a <= EPS && e <= EPS ) { return dot(r, r); }
            if( a <= EPS ) { s = 0; t = clamp01( f / e ); }
            else
            {
                const double c = dot(d1, r);
                if( e <= EPS ) { t = 0; s = clamp01( -c / a ); }
                else
                {
                    const double b = dot(d1, d2), denom = a * e - b * b;
                    s = (denom != 0) ? clamp01( (b * f - c * e) / denom ) : 0.0;
                    t = (b * s + f) / e;
                    if( t < 0 ) { t = 0; s = clamp01( -c / a ); }
                    else if( t > 1 ) { t = 1; s = clamp01( (b - c) / a ); }
                }
            }
            const V3 c1 = p1 + s * d1, c2 = p2 + t * d2;
            const V3 dd = c1 - c2;
            return dot(dd, dd);
        }

        // Does segment [p,q] cross triangle (a,b,c)? (Moller-Trumbore on the segment.)
        inline bool SegmentCrossesTriangle( V3 p, V3 q, V3 a, V3 b, V3 c )
        {
            const V3 dir = q - p, e1 = b - a, e2 = c - a;
            const V3 h = cross(dir, e2);
            const double det = dot(e1, h);
            const double scale = std::sqrt( dot(e1,e1) * dot(dir,dir) * dot(e2,e2) );
            if( std::abs(det) <= 1e-14 * scale ) { return false; }   // parallel: handled by distance queries
            const double inv = 1.0 / det;
            const V3 s = p - a;
            const double u = inv * dot(s, h);
            if( u < 0 || u > 1 ) { return false; }
            const V3 qq = cross(s, e1);
            const double v = inv * dot(dir, qq);
            if( v < 0 || u + v > 1 ) { return false; }
            const double t = inv * dot(e2, qq);
            return (t >= 0 && t <= 1);
        }
    }

    // Squared distance between simplices with `np` and `nq` points (1..3), coordinates row-major.
    inline double SimplexSquaredDistance( const double * P, const int np, const double * Q, const int nq )
    {
        using namespace detail;
        V3 p[3], q[3];
        for( int i = 0; i < np; ++i ) { p[i] = { P[3*i], P[3*i+1], P[3*i+2] }; }
        for( int i = 0; i < nq; ++i ) { q[i] = { Q[3*i], Q[3*i+1], Q[3*i+2] }; }
        double d2 = std::numeric_limits<double>::max();
        auto pt_vs = [&]( V3 x, const V3 * s, int ns ) -> double
        {
            if( ns == 1 ) { const V3 d = x - s[0]; return dot(d,d); }
            if( ns == 2 ) { return SegSegDist2( x, x, s[0], s[1] ); }
            const V3 c = ClosestPtPointTriangle( x, s[0], s[1], s[2] );
            const V3 d = x - c; return dot(d,d);
        };
        if( np == 3 && nq == 3 )
        {
            for( int i = 0; i < 3; ++i )
            {
                if( SegmentCrossesTriangle( p[i], p[(i+1)%3], q[0], q[1], q[2] ) ) { return 0.0; }
                if( SegmentCrossesTriangle( q[i], q[(i+1)%3], p[0], p[1], p[2] ) ) { return 0.0; }
            }
        }
        if( np == 2 && nq == 3 && SegmentCrossesTriangle( p[0], p[1], q[0], q[1], q[2] ) ) { return 0.0; }
        if( np == 3 && nq == 2 && SegmentCrossesTriangle( q[0], q[1], p[0], p[1], p[2] ) ) { return 0.0; }
        for( int i = 0; i < np; ++i ) { d2 = std::min( d2, pt_vs( p[i], q, nq ) ); }
        for( int i = 0; i < nq; ++i ) { d2 = std::min( d2, pt_vs( q[i], p, np ) ); }
        if( np >= 2 && nq >= 2 )
        {
            const int ep = (np == 2) ? 1 : 3, eq = (nq == 2) ? 1 : 3;
            for( int i = 0; i < ep; ++i )
            {
                for( int j = 0; j < eq; ++j )
                {
                    d2 = std::min( d2, SegSegDist2( p[i], p[(i+1)%np], q[j], q[(j+1)%nq] ) );
                }
            }
        }
        return d2;
    }

} // namespace Repulsor
