#pragma once
// Device-side tangent-point pair arithmetic (TP0 kernel), shared by the near- and far-field kernels.
//
// For the pair (S side: area a, center x, projector P; T side: b, y, Q), v = y - x:
//   phi = v.Pv, psi = v.Qv, r2 = |v|^2,
//   E   = (|phi|^{q/2-1} phi + |psi|^{q/2-1} psi) r2^{-p/2}          (q/2 >= 1)
//   g   = q r^{-p} (|phi|^{q/2-1} Pv + |psi|^{q/2-1} Qv) - p r^{-p-2} Num v
//   derivatives with respect to the S-side moments (a, a x, a P):
//     d/da  += b (E - q/2 r^{-p} |phi|^{q/2-1} phi + g.x),  d/d(a x) += -b g,  d/d(a P_k) += b/2 q r^{-p} |phi|^{q/2-1} V_k
//   metric block a1 = -2 r2^{s_exp}, a0 = (phi+psi)/(2 r2^2) a1
// (reference src/Energies/TangentPointEnergy0/TP0_Kernel_NF.hpp:140-338, TP0_Kernel_FF.hpp:130-319).
// The T-side derivative of a pair equals the S-side derivative of the swapped pair, so the device
// kernels traverse both orientations of the symmetric pattern and only accumulate S-side terms.

#include "rp_internal.hpp"

#include <type_traits>

namespace rp
{
    // Leaf of a leaf-tile workgroup. Workgroups are dealt to the 8 XCDs round robin (workgroup b runs on
    // XCD b mod 8), each with its own L2; mapping b -> (b mod 8) * per + b / 8 gives every XCD one
    // contiguous run of leaves, whose column lists overlap, so the gathered column vectors are reused
    // in that XCD's L2. The grid is padded to a multiple of 8 (launch with leaf_grid(L)); returns -1 for
    // the padding workgroups.
    __device__ __forceinline__ int64_t xcd_leaf( int64_t L )
    {
        const int64_t per = gridDim.x / NUM_XCD;
        const int64_t I = (int64_t)( blockIdx.x % NUM_XCD ) * per + blockIdx.x / NUM_XCD;
        return ( I < L ) ? I : -1;
    }
    inline dim3 leaf_grid( int64_t L ) { return dim3( (unsigned)( ( (L + NUM_XCD - 1) / NUM_XCD ) * NUM_XCD ) ); }

    // Exponent kinds compiled as fast paths; KIND_GENERIC reads the exponents at run time.
    enum { KIND_GENERIC = 0, KIND_Q6P12 = 1, KIND_Q2P4 = 2 };

    template<typename R>
    __device__ __forceinline__ R ipow( R x, int n )   // n >= 0
    {
        R r = R(1), b = x;
        while( n > 0 ) { if( n & 1 ) { r *= b; } b *= b; n >>= 1; }
        return r;
    }

    __device__ __forceinline__ double dev_cbrt( double x ) { return cbrt(x); }
    __device__ __forceinline__ float  dev_cbrt( float x )  { return cbrtf(x); }
    __device__ __forceinline__ double dev_pow( double x, double y ) { return pow(x, y); }
    __device__ __forceinline__ float  dev_pow( float x, float y )   { return powf(x, y); }

    // r2^{s_exp} for the metric kernel.
    template<typename R>
    __device__ __forceinline__ R metric_power( const Expo & ex, R r2 )
    {
        if( ex.s_mode == 0 )
        {
            const int k = ex.s_k;
            return (k >= 0) ? ipow( r2, k ) : R(1) / ipow( r2, -k );
        }
        else if( ex.s_mode == 1 )
        {
            // r2^{k/3} = r2^{floor(k/3)} * cbrt(r2)^{k - 3 floor(k/3)}
            const int k = ex.s_k;
            const int f = (k >= 0) ? (k / 3) : -((-k + 2) / 3);
            const int rem = k - 3 * f;           // 1 or 2
            const R c = dev_cbrt( r2 );
            const R cr = (rem == 1) ? c : c * c;
            const R pf = (f >= 0) ? ipow( r2, f ) : R(1) / ipow( r2, -f );
            return pf * cr;
        }
        return dev_pow( r2, (R)ex.s_exp );
    }

    template<typename R>
    struct PairTerms
    {
        R E;        // energy density (without areas)
        R E0;       // E - q/2 r^{-p} |phi|^{q/2-1} phi
        R g[3];     // dE/dv
        R hK;       // q/2 r^{-p} |phi|^{q/2-1}
        R Vk[6];    // (2 - delta_ab) v_a v_b packed
        R a0, a1;   // metric block
    };

    // PROJ: P and Q are orthogonal projectors (primitive rows; cluster rows hold area-weighted averages of
    // projectors, which are not).
    template<typename R, int KIND, bool NEED_E, bool NEED_D, bool NEED_M, bool PROJ = false>
    __device__ __forceinline__ void pair_terms( const Expo & ex, const R v[3], const R P[6], const R Q[6], PairTerms<R> & t )
    {
        const R Pv0 = P[0] * v[0] + P[1] * v[1] + P[2] * v[2];
        const R Pv1 = P[1] * v[0] + P[3] * v[1] + P[4] * v[2];
        const R Pv2 = P[2] * v[0] + P[4] * v[1] + P[5] * v[2];
        const R Qv0 = Q[0] * v[0] + Q[1] * v[1] + Q[2] * v[2];
        const R Qv1 = Q[1] * v[0] + Q[3] * v[1] + Q[4] * v[2];
        const R Qv2 = Q[2] * v[0] + Q[4] * v[1] + Q[5] * v[2];
        const R r2  = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        // phi = <v, P v>. For an orthogonal projector (P = P^T P) <v, P v> = |P v|^2. FP64 keeps the
        // reference's form; FP32 on projectors uses |P v|^2: for close pairs v lies almost in the tangent
        // space, and <v, P v> then sums terms of size |v|^2 to a result of size |v|^2 theta^2 (relative
        // rounding about eps / theta^2), while |P v|^2 squares components of size |v| theta (eps / theta).
        R phi, psi;
        if constexpr ( PROJ && std::is_same<R, float>::value )
        {
            phi = Pv0 * Pv0 + Pv1 * Pv1 + Pv2 * Pv2;
            psi = Qv0 * Qv0 + Qv1 * Qv1 + Qv2 * Qv2;
        }
        else
        {
            phi = v[0] * Pv0 + v[1] * Pv1 + v[2] * Pv2;
            psi = v[0] * Qv0 + v[1] * Qv1 + v[2] * Qv2;
        }

        if constexpr ( NEED_E || NEED_D )
        {
            R rmp2;   // r2^{-p/2-1}
            R phq2, psq2, phq, psq;
            if constexpr ( KIND == KIND_Q6P12 )
            {
                // r2^{-7}: one reciprocal of r2^7 = r2^4 * r2^3
                const R r2_2 = r2 * r2;
                const R r2_3 = r2_2 * r2;
                const R r2_4 = r2_2 * r2_2;
                rmp2 = R(1) / ( r2_4 * r2_3 );
                phq2 = phi * phi;
                psq2 = psi * psi;
                phq  = phq2 * phi;
                psq  = psq2 * psi;
            }
            else if constexpr ( KIND == KIND_Q2P4 )
            {
                rmp2 = R(1) / ( r2 * r2 * r2 );     // r2^{-3}
                phq2 = R(1);
                psq2 = R(1);
                phq  = phi;
                psq  = psi;
            }
            else
            {
                rmp2 = ex.p_int ? R(1) / ipow( r2, ex.p_half_i + 1 ) : dev_pow( r2, (R)(-ex.p_half - 1.0) );
                if( ex.q_flag >= 2 )
                {
                    phq2 = ex.q_int ? ipow( fabs(phi), ex.q_half_i - 1 ) : dev_pow( (R)fabs(phi), (R)(ex.q_half - 1.0) );
                    psq2 = ex.q_int ? ipow( fabs(psi), ex.q_half_i - 1 ) : dev_pow( (R)fabs(psi), (R)(ex.q_half - 1.0) );
                    phq  = phq2 * phi;
                    psq  = psq2 * psi;
                }
                else
                {
                    phq2 = R(1);
                    psq2 = R(1);
                    phq  = ex.q_int ? ipow( fabs(phi), ex.q_half_i ) : dev_pow( (R)fabs(phi), (R)ex.q_half );
                    psq  = ex.q_int ? ipow( fabs(psi), ex.q_half_i ) : dev_pow( (R)fabs(psi), (R)ex.q_half );
                }
            }
            const R rmp = rmp2 * r2;
            const R Num = phq + psq;
            t.E = Num * rmp;
            if constexpr ( NEED_D )
            {
                const bool qz = ( KIND == KIND_GENERIC ) && ( ex.q_flag == 0 );
                const R qv = ( KIND == KIND_Q6P12 ) ? R(6) : ( ( KIND == KIND_Q2P4 ) ? R(2) : (R)ex.q );
                const R pv = ( KIND == KIND_Q6P12 ) ? R(12) : ( ( KIND == KIND_Q2P4 ) ? R(4) : (R)ex.p );
                const R factor = qz ? R(0) : qv * rmp;
                const R Kxy = qz ? R(0) : factor * phq2;
                const R Kyx = qz ? R(0) : factor * psq2;
                const R H = pv * rmp2 * Num;
                t.g[0] = ( qz ? R(0) : Kxy * Pv0 + Kyx * Qv0 ) - H * v[0];
                t.g[1] = ( qz ? R(0) : Kxy * Pv1 + Kyx * Qv1 ) - H * v[1];
                t.g[2] = ( qz ? R(0) : Kxy * Pv2 + Kyx * Qv2 ) - H * v[2];
                t.E0 = t.E - R(0.5) * factor * phq;
                t.hK = R(0.5) * Kxy;
                t.Vk[0] = v[0] * v[0];
                t.Vk[1] = R(2) * v[0] * v[1];
                t.Vk[2] = R(2) * v[0] * v[2];
                t.Vk[3] = v[1] * v[1];
                t.Vk[4] = R(2) * v[1] * v[2];
                t.Vk[5] = v[2] * v[2];
            }
        }
        if constexpr ( NEED_M )
        {
            const R r4 = r2 * r2;
            t.a1 = R(-2) * metric_power<R>( ex, r2 );
            t.a0 = R(0.5) * ( phi + psi ) / r4 * t.a1;
        }
    }

} // namespace rp
