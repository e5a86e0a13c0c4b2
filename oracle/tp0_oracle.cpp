// ORACLE -- test infrastructure only.
//
// CPU restatement (FP64, single thread, reference summation order) of the reference's
// tangent-point hot path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load this library, and only as the checker or the timed CPU baseline; the product path never
// calls it.
//
// Restated reference functions (file:line in HenrikSchumacher/Repulsor):
//   orc_simplex_data          SimplexDataKernel::ReadSimplex/WriteNear/WriteFar/WriteDiffOp
//                             src/SimplexDataKernel.hpp:92-238
//   orc_tp0_compute           TP0_Traversor::Compute / NF_Compute / FF_Compute
//                             src/Energies/TangentPointEnergy0/TP0_Traversor.hpp:129-376,
//                             FMM_Traversor::Compute src/FMM/FMM_Traversor.hpp:56-147,
//                             FMM_Kernel_NF::loadS/loadT/writeS/writeT src/FMM/FMM_Kernel_NF.hpp:169-245,
//                             FMM_Kernel_FF::loadS/loadT/writeS/writeT src/FMM/FMM_Kernel_FF.hpp:162-217,
//                             TP0_Kernel_NF::Compute src/Energies/TangentPointEnergy0/TP0_Kernel_NF.hpp:140-338,
//                             TP0_Kernel_FF::Compute src/Energies/TangentPointEnergy0/TP0_Kernel_FF.hpp:130-319
//   orc_collect_derivatives   ClusterTree::CollectNearFieldDerivatives / CollectFarFieldDerivatives
//                             src/ClusterTree.hpp:1115-1208, PercolateDown src/ClusterTree/Percolate_Parallel.hpp:187-287
//   orc_dnearfar_to_hulls     SimplicialMeshDetails<2,3>::DNearFarToHulls src/SimplicialMesh/SimplicialMeshDetails.hpp:1097-1410
//                             and <1,3> :529-736 (spec CodeGeneration/SimplicialMesh/Code_DNearFar_to_Hulls.m),
//                             re-derived by hand from the moment map (a, a*hull, a*P) / (a, a*x, a*P).
//   orc_assemble              DerivativeAssembler().Dot<3> src/SimplicialMesh.hpp:625-675
//   orc_multiply_metric       MetricDimRestricted::MultiplyMetric src/Metrics/MetricDimRestricted.hpp:109-139,
//                             ClusterTree::Pre/Post src/ClusterTree.hpp:770-943,
//                             TP0_Traversor::MultiplyMetric src/Energies/TangentPointEnergy0/TP0_Traversor.hpp:384-505,
//                             TP0_Kernel_MultiplyMetric_Tiny::ApplyBlock .../TP0_Kernel_MultiplyMetric_Tiny.hpp:89-138
//   orc_allpairs_energy       TangentPointEnergy_AllPairs idea (src/Energies/TangentPointEnergy/TP_AllPairs_Traversor.hpp):
//                             sum over all unordered pairs of a_i a_j E_ij, no tree.
//   orc_pl_compute            PseudoLaplacian_Traversor::Compute (LowOrder)
//                             src/Metrics/PseudoLaplacian/PseudoLaplacian_Traversor.hpp:143-300,
//                             PseudoLaplacian_Kernel_NF.hpp:124-146, PseudoLaplacian_Kernel_FF.hpp:101-146
//   orc_pl_multiply           MetricDimRestricted::MultiplyMetric with OperatorType::LowOrder
//                             src/Metrics/MetricDimRestricted.hpp:109-139, ClusterTree::Pre/Post (lo_pre, lo_post)
//                             src/ClusterTree.hpp:458-500,770-943
//
// The *_mt variants take a thread count: one thread runs the rows in order (what the checker uses), more
// threads split the rows by cost with per-thread accumulators, as the reference's ParallelDoReduce /
// AddReduce do (FMM_Traversor.hpp:56-147); they serve as the multi-core CPU baseline.
//
// The `Power` of the Tensors library (absent from the reference checkout) is restated as binary
// powering for integer exponents and std::pow for real exponents.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

typedef int32_t Int;
typedef int64_t LInt;

namespace
{
    struct Neumaier
    {
        double s = 0, c = 0;
        void add( double x )
        {
            const double t = s + x;
            if( std::abs(s) >= std::abs(x) ) { c += (s - t) + x; } else { c += (x - t) + s; }
            s = t;
        }
        double value() const { return s + c; }
    };

    double IntPow( double x, LInt n )
    {
        if( n < 0 ) { return 1.0 / IntPow( x, -n ); }
        double r = 1.0, b = x;
        while( n > 0 ) { if( n & 1 ) { r *= b; } b *= b; n >>= 1; }
        return r;
    }

    int lin_k3( int i, int j ) { if( i > j ) { int t = i; i = j; j = t; } return i * 3 - (i * (i - 1)) / 2 + (j - i); }

    // Static partition of the rows [0, rows) into `parts` contiguous ranges of about equal cost (cost(i) + 1
    // per row): the reference's JobPointers over the rows' pair counts (FMM_Traversor.hpp:65-69).
    template<typename Cost>
    std::vector<LInt> CostPartition( const LInt rows, const int parts, Cost && cost )
    {
        std::vector<double> pre( (size_t)rows + 1, 0.0 );
        for( LInt i = 0; i < rows; ++i ) { pre[(size_t)i + 1] = pre[(size_t)i] + (double)cost( i ) + 1.0; }
        std::vector<LInt> cut( (size_t)parts + 1, rows );
        cut[0] = 0;
        for( int t = 1; t < parts; ++t )
        {
            const double target = pre[(size_t)rows] * t / parts;
            const LInt k = (LInt)( std::lower_bound( pre.begin(), pre.end(), target ) - pre.begin() );
            cut[(size_t)t] = std::max( cut[(size_t)t - 1], std::min( rows, k ) );
        }
        return cut;
    }

    // Runs f(t, begin, end) for the ranges of cut, one std::thread per range (the caller's thread alone
    // for a single range).
    template<typename F>
    void ParallelRanges( const std::vector<LInt> & cut, F && f )
    {
        const int parts = (int)cut.size() - 1;
        if( parts <= 1 ) { f( 0, cut[0], cut[1] ); return; }
        std::vector<std::thread> pool;
        pool.reserve( (size_t)parts );
        for( int t = 0; t < parts; ++t ) { pool.emplace_back( [&f, &cut, t]() { f( t, cut[(size_t)t], cut[(size_t)t + 1] ); } ); }
        for( auto & th : pool ) { th.join(); }
    }

    // out[0..n) += sum over the per-thread buffers bufs[1..] (thread order), split over threads.
    void ReduceBuffers( double * out, const std::vector<std::vector<double>> & bufs, const LInt n, const int threads )
    {
        if( bufs.size() <= 1 || n <= 0 ) { return; }
        std::vector<LInt> cut( (size_t)threads + 1 );
        for( int t = 0; t <= threads; ++t ) { cut[(size_t)t] = n * t / threads; }
        ParallelRanges( cut, [&]( int, LInt b, LInt e )
        {
            for( size_t t = 1; t < bufs.size(); ++t ) { for( LInt k = b; k < e; ++k ) { out[k] += bufs[t][(size_t)k]; } }
        });
    }

    // Exponent specialisation of TP0_Traversor::Compute (TP0_Traversor.hpp:169-245).
    struct Exponents
    {
        double q, p, q_half, p_half;
        bool q_int, p_int;
        LInt q_half_i, p_half_i;
        int q_flag;    // 2, 0, or -1 (invalid)
        double s, s_exp;

        Exponents( double q_, double p_, int n )
        {
            q = q_; p = p_;
            q_half = q / 2; p_half = p / 2;
            q_half_i = (LInt)q_half; p_half_i = (LInt)p_half;
            q_int = (q_half == (double)q_half_i);
            p_int = (p_half == (double)p_half_i);
            q_flag = (q_half >= 1) ? 2 : ((q_half == 0) ? 0 : -1);
            s = (p - n) / q;
            s_exp = -0.5 * (2.0 * (s - 1.0) + n);
        }
        double PowQ( double x, double e_real, LInt e_int ) const { return q_int ? IntPow(x, e_int) : std::pow(x, e_real); }
        double PowP( double x, double e_real, LInt e_int ) const { return p_int ? IntPow(x, e_int) : std::pow(x, e_real); }
    };

    // One pair evaluation of TP0_Kernel_{NF,FF}::Compute. x,P: S side; y,Q: T side.
    // DX/DY layout: [0] area part, then `ncoord` coordinate slots (near: hull vertices, far: center),
    // then 6 projector slots. Returns the energy term (without symmetry factor, which is 1).
    struct PairOut { double result, a0, a1; };

    PairOut PairCompute( const Exponents & ex, int n, bool energy, bool diff, bool metric,
        double a, const double * x, const double * P, double b, const double * y, const double * Q,
        double * DX, double * DY, int n_hull_S, int n_hull_T, double S_scale, double T_scale )
    {
        double v[3], Pv[3] = {0,0,0}, Qv[3] = {0,0,0}, V[6] = {0,0,0,0,0,0};
        double r2 = 0, rCosPhi_2 = 0, rCosPsi_2 = 0;
        for( int i = 0; i < 3; ++i ) { v[i] = y[i] - x[i]; r2 += v[i] * v[i]; }
        for( int i = 0; i < 3; ++i )
        {
            for( int j = 0; j < 3; ++j )
            {
                const int k = lin_k3(i,j);
                Pv[i] += P[k] * v[j];
                Qv[i] += Q[k] * v[j];
                if( j >= i ) { V[k] = (1.0 + (double)(i != j)) * v[i] * v[j]; }
            }
            rCosPhi_2 += v[i] * Pv[i];
            rCosPsi_2 += v[i] * Qv[i];
        }
        PairOut out{0,0,0};
        if( energy || diff )
        {
            const double r_minus_p_minus_2 = ex.PowP( r2, -ex.p_half - 1, -ex.p_half_i - 1 );
            const double r_minus_p = r_minus_p_minus_2 * r2;
            double rCosPhi_q_minus_2, rCosPsi_q_minus_2, rCosPhi_q, rCosPsi_q;
            if( ex.q_flag >= 2 )
            {
                rCosPhi_q_minus_2 = ex.PowQ( std::abs(rCosPhi_2), ex.q_half - 1, ex.q_half_i - 1 );
                rCosPsi_q_minus_2 = ex.PowQ( std::abs(rCosPsi_2), ex.q_half - 1, ex.q_half_i - 1 );
                rCosPhi_q = rCosPhi_q_minus_2 * rCosPhi_2;
                rCosPsi_q = rCosPsi_q_minus_2 * rCosPsi_2;
            }
            else
            {
                rCosPhi_q_minus_2 = 1;
                rCosPsi_q_minus_2 = 1;
                rCosPhi_q = ex.PowQ( std::abs(rCosPhi_2), ex.q_half, ex.q_half_i );
                rCosPsi_q = ex.PowQ( std::abs(rCosPsi_2), ex.q_half, ex.q_half_i );
            }
            const double Num = rCosPhi_q + rCosPsi_q;
            const double E = Num * r_minus_p;
            if( energy ) { out.result = a * E * b; }
            if( diff )
            {
                const double factor = (ex.q_flag == 0) ? 0.0 : ex.q * r_minus_p;
                const double K_xy   = (ex.q_flag == 0) ? 0.0 : factor * rCosPhi_q_minus_2;
                const double K_yx   = (ex.q_flag == 0) ? 0.0 : factor * rCosPsi_q_minus_2;
                const double H      = ex.p * r_minus_p_minus_2 * Num;
                double dEdvx = 0, dEdvy = 0;
                for( int i = 0; i < 3; ++i )
                {
                    const double dEdv = ((ex.q_flag == 0) ? 0.0 : K_xy * Pv[i] + K_yx * Qv[i]) - H * v[i];
                    dEdvx += dEdv * x[i];
                    dEdvy += dEdv * y[i];
                    for( int ii = 0; ii < n_hull_S; ++ii ) { DX[1 + 3 * ii + i] -= S_scale * b * dEdv; }
                    for( int ii = 0; ii < n_hull_T; ++ii ) { DY[1 + 3 * ii + i] += T_scale * a * dEdv; }
                }
                DX[0] += b * ( E - 0.5 * factor * rCosPhi_q + dEdvx );
                DY[0] += a * ( E - 0.5 * factor * rCosPsi_q - dEdvy );
                const double half_b_K_xy = 0.5 * b * K_xy;
                const double half_a_K_yx = 0.5 * a * K_yx;
                for( int k = 0; k < 6; ++k )
                {
                    DX[1 + 3 * n_hull_S + k] += half_b_K_xy * V[k];
                    DY[1 + 3 * n_hull_T + k] += half_a_K_yx * V[k];
                }
            }
        }
        if( metric )
        {
            const double r4 = r2 * r2;
            const double a_1 = -2.0 * std::pow( r2, ex.s_exp );
            const double a_0 = 0.5 * (rCosPhi_2 + rCosPsi_2) / r4 * a_1;
            out.a0 = a_0; out.a1 = a_1;
        }
        (void)n;
        return out;
    }
}

extern "C"
{

// SimplexDataKernel restatement for all simplices (original order).
// near: T x (1 + 3(n+1) + 6), far: T x 10, diffop: T x 3 x (n+1) (columns = sorted simplex),
// sorted: T x (n+1).
void orc_simplex_data( int n, LInt T, const double * V, const Int * S, const double * charges,
                       double * near, double * far, double * diffop, Int * sorted )
{
    const int SIZE = n + 1;
    const int ND = 1 + 3 * SIZE + 6;
    for( LInt t = 0; t < T; ++t )
    {
        Int s[3], ss[3];
        for( int l = 0; l < SIZE; ++l ) { s[l] = S[SIZE * t + l]; ss[l] = s[l]; }
        std::sort( ss, ss + SIZE );
        double hull[3][3], center[3];
        for( int l = 0; l < SIZE; ++l ) { for( int k = 0; k < 3; ++k ) { hull[l][k] = V[3 * (LInt)s[l] + k]; } }
        double charge = charges ? charges[s[0]] : 1.0;
        for( int k = 0; k < 3; ++k ) { center[k] = hull[0][k]; }
        for( int l = 1; l < SIZE; ++l ) { for( int k = 0; k < 3; ++k ) { center[k] += hull[l][k]; } charge += charges ? charges[s[l]] : 1.0; }
        for( int k = 0; k < 3; ++k ) { center[k] *= 1.0 / SIZE; }
        charge *= 1.0 / SIZE;
        double df[3][2];
        for( int l = 0; l < n; ++l ) { for( int k = 0; k < 3; ++k ) { df[k][l] = V[3 * (LInt)ss[l+1] + k] - V[3 * (LInt)ss[0] + k]; } }
        double g[2][2];
        for( int r = 0; r < n; ++r ) { for( int c = 0; c < n; ++c ) { double acc = 0; for( int k = 0; k < 3; ++k ) { acc += df[k][r] * df[k][c]; } g[r][c] = acc; } }
        double a = charge * (n == 2 ? 0.5 : 1.0);
        double ginv[2][2];
        if( n == 2 )
        {
            const double det = g[0][0] * g[1][1] - g[0][1] * g[1][0];
            a *= std::sqrt(det);
            ginv[0][0] =  g[1][1] / det; ginv[1][1] = g[0][0] / det;
            ginv[0][1] = -g[0][1] / det; ginv[1][0] = -g[1][0] / det;
        }
        else
        {
            a *= std::sqrt(g[0][0]);
            ginv[0][0] = 1.0 / g[0][0];
        }
        double dfd[2][3];
        for( int l = 0; l < n; ++l ) { for( int k = 0; k < 3; ++k ) { double acc = 0; for( int m = 0; m < n; ++m ) { acc += ginv[l][m] * df[k][m]; } dfd[l][k] = acc; } }
        double P[3][3];
        for( int r = 0; r < 3; ++r ) { for( int c = 0; c < 3; ++c ) { double acc = 0; for( int l = 0; l < n; ++l ) { acc += df[r][l] * dfd[l][c]; } P[r][c] = (r == c ? 1.0 : 0.0) - acc; } }
        double * nr = near + ND * t;
        nr[0] = a;
        for( int l = 0; l < SIZE; ++l ) { for( int k = 0; k < 3; ++k ) { nr[1 + 3 * l + k] = hull[l][k]; } }
        const int pi[6] = {0,0,0,1,1,2}, pj[6] = {0,1,2,1,2,2};
        for( int k = 0; k < 6; ++k ) { nr[1 + 3 * SIZE + k] = P[pi[k]][pj[k]]; }
        double * fr = far + 10 * t;
        fr[0] = a;
        for( int k = 0; k < 3; ++k ) { fr[1 + k] = center[k]; }
        for( int k = 0; k < 6; ++k ) { fr[4 + k] = P[pi[k]][pj[k]]; }
        if( diffop )
        {
            double * D = diffop + 3 * SIZE * t;
            for( int k = 0; k < 3; ++k )
            {
                D[SIZE * k] = 0;
                for( int l = 0; l < n; ++l ) { D[SIZE * k] -= dfd[l][k]; D[SIZE * k + l + 1] = dfd[l][k]; }
            }
        }
        if( sorted ) { for( int l = 0; l < SIZE; ++l ) { sorted[SIZE * t + l] = ss[l]; } }
    }
}

// TP0_Traversor::Compute for a symmetric BCT (S == T). Patterns are full symmetric CSR with
// `diag` = first strictly-upper position per row. flags: 1 energy, 2 differential, 4 metric.
// P_near: T x ND (tree order), C_far: C x 10.
// Outputs (any may be NULL when not requested): P_D_near T x ND, C_D_far C x 10 (both zeroed here),
// nf_vals N_nf x 2, ff_vals N_ff x 2 (full pattern, lower part filled from the upper part),
// nf_diag T x 2, ff_diag C x 2. Returns the energy (compensated sum).
// threads = 1 runs the rows in order (the checker); more threads split the rows of each pattern by
// cost, with per-thread energy sums, derivative and diagonal buffers added in thread order at the end
// (the reference's per-thread buffers and AddReduce, TP0_Traversor.hpp:145-167, 306, 340).
double orc_tp0_compute_mt( int n, LInt T, LInt C, const double * P_near, const double * C_far,
    const LInt * nrp, const Int * nci, const LInt * ndiag,
    const LInt * frp, const Int * fci, const LInt * fdiag,
    double q, double p, int flags,
    double * P_D_near, double * C_D_far, double * nf_vals, double * ff_vals, double * nf_diag, double * ff_diag, int threads )
{
    const bool energy = flags & 1, diff = flags & 2, metric = flags & 4;
    const Exponents ex( q, p, n );
    if( ex.q_flag < 0 ) { return std::nan(""); }
    threads = std::max( 1, threads );
    const int SIZE = n + 1;
    const int ND = 1 + 3 * SIZE + 6;
    if( diff )
    {
        std::memset( P_D_near, 0, sizeof(double) * (size_t)(ND * T) );
        std::memset( C_D_far,  0, sizeof(double) * (size_t)(10 * C) );
    }
    if( metric )
    {
        std::memset( nf_diag, 0, sizeof(double) * (size_t)(2 * T) );
        std::memset( ff_diag, 0, sizeof(double) * (size_t)(2 * C) );
    }
    const double scale = 1.0 / SIZE;
    std::vector<Neumaier> en( (size_t)threads );
    // per-thread derivative / diagonal buffers: thread 0 writes the outputs, the others their own
    auto buffers = [&]( const bool need ) { return std::vector<std::vector<double>>( need ? (size_t)threads : 0 ); };
    // ---- near field ----
    if( nrp[T] > 0 )
    {
        const std::vector<LInt> cut = CostPartition( T, threads, [&]( LInt i ) { return nrp[i + 1] - ndiag[i]; } );
        auto dbuf = buffers( diff ), gbuf = buffers( metric );
        ParallelRanges( cut, [&]( int t, LInt i0, LInt i1 )
        {
            double * PD = P_D_near;
            double * DG = nf_diag;
            if( t > 0 && diff )   { dbuf[(size_t)t].assign( (size_t)(ND * T), 0.0 ); PD = dbuf[(size_t)t].data(); }
            if( t > 0 && metric ) { gbuf[(size_t)t].assign( (size_t)(2 * T), 0.0 ); DG = gbuf[(size_t)t].data(); }
            Neumaier e = en[(size_t)t];     // thread-local accumulator, stored once (no false sharing)
            for( LInt i = i0; i < i1; ++i )
            {
                const LInt kb = ndiag[i], ke = nrp[i + 1];
                if( ke <= kb ) { continue; }
                const double * X = P_near + ND * i;
                const double a = X[0];
                double x[3];
                for( int k = 0; k < 3; ++k ) { x[k] = scale * X[1 + k]; for( int kk = 1; kk < SIZE; ++kk ) { x[k] += scale * X[1 + 3 * kk + k]; } }
                double DX[16] = {0};
                double ii_block[2] = {0,0};
                for( LInt k = kb; k < ke; ++k )
                {
                    const Int j = nci[k];
                    const double * Y = P_near + ND * j;
                    const double b = Y[0];
                    double y[3];
                    for( int kk = 0; kk < 3; ++kk ) { y[kk] = scale * Y[1 + kk]; for( int h = 1; h < SIZE; ++h ) { y[kk] += scale * Y[1 + 3 * h + kk]; } }
                    double DY[16] = {0};
                    const PairOut o = PairCompute( ex, n, energy, diff, metric, a, x, X + 1 + 3 * SIZE, b, y, Y + 1 + 3 * SIZE,
                                                   DX, DY, SIZE, SIZE, scale, scale );
                    if( energy ) { e.add( o.result ); }
                    if( diff ) { for( int h = 0; h < ND; ++h ) { PD[ND * j + h] += DY[h]; } }
                    if( metric )
                    {
                        nf_vals[2 * k] = o.a0; nf_vals[2 * k + 1] = o.a1;
                        ii_block[0] -= (b / a) * o.a0;
                        ii_block[1] -= (b / a) * o.a1;
                        DG[2 * j]     += -(a / b) * o.a0;
                        DG[2 * j + 1] += -(a / b) * o.a1;
                    }
                }
                if( diff ) { for( int h = 0; h < ND; ++h ) { PD[ND * i + h] += DX[h]; } }
                if( metric ) { DG[2 * i] += ii_block[0]; DG[2 * i + 1] += ii_block[1]; }
            }
            en[(size_t)t] = e;
        });
        if( diff ) { ReduceBuffers( P_D_near, dbuf, ND * T, threads ); }
        if( metric )
        {
            ReduceBuffers( nf_diag, gbuf, 2 * T, threads );
            // FillLowerTriangleFromUpperTriangle.
            ParallelRanges( CostPartition( T, threads, [&]( LInt i ) { return ndiag[i] - nrp[i]; } ), [&]( int, LInt i0, LInt i1 )
            {
                for( LInt i = i0; i < i1; ++i )
                {
                    for( LInt k = nrp[i]; k < ndiag[i]; ++k )
                    {
                        const Int j = nci[k];
                        const Int * b = nci + ndiag[j];
                        const Int * e = nci + nrp[j + 1];
                        const Int * f = std::lower_bound( b, e, (Int)i );
                        const LInt pos = f - nci;
                        nf_vals[2 * k] = nf_vals[2 * pos]; nf_vals[2 * k + 1] = nf_vals[2 * pos + 1];
                    }
                }
            });
        }
    }
    // ---- far field ----
    if( frp[C] > 0 )
    {
        const std::vector<LInt> cut = CostPartition( C, threads, [&]( LInt i ) { return frp[i + 1] - fdiag[i]; } );
        auto dbuf = buffers( diff ), gbuf = buffers( metric );
        ParallelRanges( cut, [&]( int t, LInt i0, LInt i1 )
        {
            double * CD = C_D_far;
            double * DG = ff_diag;
            if( t > 0 && diff )   { dbuf[(size_t)t].assign( (size_t)(10 * C), 0.0 ); CD = dbuf[(size_t)t].data(); }
            if( t > 0 && metric ) { gbuf[(size_t)t].assign( (size_t)(2 * C), 0.0 ); DG = gbuf[(size_t)t].data(); }
            Neumaier e = en[(size_t)t];
            for( LInt i = i0; i < i1; ++i )
            {
                const LInt kb = fdiag[i], ke = frp[i + 1];
                if( ke <= kb ) { continue; }
                const double * X = C_far + 10 * i;
                const double a = X[0];
                double DX[10] = {0};
                double ii_block[2] = {0,0};
                for( LInt k = kb; k < ke; ++k )
                {
                    const Int j = fci[k];
                    const double * Y = C_far + 10 * j;
                    const double b = Y[0];
                    double DY[10] = {0};
                    const PairOut o = PairCompute( ex, n, energy, diff, metric, a, X + 1, X + 4, b, Y + 1, Y + 4, DX, DY, 1, 1, 1.0, 1.0 );
                    if( energy ) { e.add( o.result ); }
                    if( diff ) { for( int h = 0; h < 10; ++h ) { CD[10 * j + h] += DY[h]; } }
                    if( metric )
                    {
                        ff_vals[2 * k] = o.a0; ff_vals[2 * k + 1] = o.a1;
                        ii_block[0] -= (b / a) * o.a0;
                        ii_block[1] -= (b / a) * o.a1;
                        DG[2 * j]     += -(a / b) * o.a0;
                        DG[2 * j + 1] += -(a / b) * o.a1;
                    }
                }
                if( diff ) { for( int h = 0; h < 10; ++h ) { CD[10 * i + h] += DX[h]; } }
                if( metric ) { DG[2 * i] += ii_block[0]; DG[2 * i + 1] += ii_block[1]; }
            }
            en[(size_t)t] = e;
        });
        if( diff ) { ReduceBuffers( C_D_far, dbuf, 10 * C, threads ); }
        if( metric )
        {
            ReduceBuffers( ff_diag, gbuf, 2 * C, threads );
            for( LInt i = 0; i < C; ++i )
            {
                for( LInt k = frp[i]; k < fdiag[i]; ++k )
                {
                    const Int j = fci[k];
                    const Int * f = std::lower_bound( fci + fdiag[j], fci + frp[j + 1], (Int)i );
                    const LInt pos = f - fci;
                    ff_vals[2 * k] = ff_vals[2 * pos]; ff_vals[2 * k + 1] = ff_vals[2 * pos + 1];
                }
            }
        }
    }
    Neumaier total;
    for( const auto & e : en ) { total.add( e.s ); total.add( e.c ); }
    return threads == 1 ? en[0].value() : total.value();
}

double orc_tp0_compute( int n, LInt T, LInt C, const double * P_near, const double * C_far,
    const LInt * nrp, const Int * nci, const LInt * ndiag,
    const LInt * frp, const Int * fci, const LInt * fdiag,
    double q, double p, int flags,
    double * P_D_near, double * C_D_far, double * nf_vals, double * ff_vals, double * nf_diag, double * ff_diag )
{
    return orc_tp0_compute_mt( n, T, C, P_near, C_far, nrp, nci, ndiag, frp, fci, fdiag, q, p, flags,
                               P_D_near, C_D_far, nf_vals, ff_vals, nf_diag, ff_diag, 1 );
}

// CollectFarFieldDerivatives + CollectNearFieldDerivatives. Inputs in tree order; outputs in the
// original primitive order. C_D_far is used as scratch (percolated down in place).
void orc_collect_derivatives( int n, LInt T, LInt C, const Int * C_left, const Int * C_right,
    const Int * C_begin, const Int * C_end, const Int * P_inverse_ordering,
    const double * P_D_near_tree, double * C_D_far, double * P_D_near, double * P_D_far )
{
    const int ND = 1 + 3 * (n + 1) + 6;
    // PercolateDown: parents before children in depth-first numbering.
    for( LInt c = 0; c < C; ++c )
    {
        const Int L = C_left[c];
        if( L >= 0 )
        {
            const Int R = C_right[c];
            for( int h = 0; h < 10; ++h ) { C_D_far[10 * L + h] = C_D_far[10 * c + h] + C_D_far[10 * L + h]; }
            for( int h = 0; h < 10; ++h ) { C_D_far[10 * R + h] = C_D_far[10 * c + h] + C_D_far[10 * R + h]; }
        }
    }
    std::vector<double> P_out( (size_t)(10 * T), 0.0 );
    for( LInt c = 0; c < C; ++c )
    {
        if( C_left[c] < 0 )
        {
            for( Int i = C_begin[c]; i < C_end[c]; ++i ) { for( int h = 0; h < 10; ++h ) { P_out[10 * (size_t)i + h] = C_D_far[10 * c + h]; } }
        }
    }
    for( LInt i = 0; i < T; ++i )
    {
        const LInt from = P_inverse_ordering[i];
        for( int h = 0; h < 10; ++h ) { P_D_far[10 * i + h] = P_out[10 * from + h]; }
        for( int h = 0; h < ND; ++h ) { P_D_near[ND * i + h] = P_D_near_tree[ND * from + h]; }
    }
}

// Chain rule from moment derivatives to vertex coordinates (DNearFarToHulls).
// N: T x ND near derivatives, F: T x 10 far derivatives (original order); B: T x (n+1) x 3.
// Moment map per simplex with vertices X_0..X_n (input order):
//   triangles: nu = (X_1-X_0) x (X_2-X_0), vol = |nu|/2, pr = nu nu^T / |nu|^2
//   edges:     tau = X_1-X_0,              vol = |tau|,  pr = I - tau tau^T / |tau|^2
//   near moments (vol, vol*hull, vol*pr_packed), far moments (vol, vol*mean(hull), vol*pr_packed);
//   B = charge * ( N . d(near moments)/dX + F . d(far moments)/dX ).
void orc_dnearfar_to_hulls( int n, LInt T, const double * V, const Int * S, const double * charges,
    const double * N, const double * F, double * B )
{
    const int SIZE = n + 1;
    const int ND = 1 + 3 * SIZE + 6;
    const int pi[6] = {0,0,0,1,1,2}, pj[6] = {0,1,2,1,2,2};
    for( LInt t = 0; t < T; ++t )
    {
        double X[3][3];
        double charge = 0;
        for( int l = 0; l < SIZE; ++l ) { for( int k = 0; k < 3; ++k ) { X[l][k] = V[3 * (LInt)S[SIZE * t + l] + k]; } charge += charges ? charges[S[SIZE * t + l]] : 1.0; }
        charge *= 1.0 / SIZE;
        const double * Nt = N + ND * t;
        const double * Ft = F + 10 * t;
        double vol, pr[3][3], w[3];   // w: nu (triangles) or tau (edges)
        double ww;
        if( n == 2 )
        {
            const double e1[3] = { X[1][0]-X[0][0], X[1][1]-X[0][1], X[1][2]-X[0][2] };
            const double e2[3] = { X[2][0]-X[0][0], X[2][1]-X[0][1], X[2][2]-X[0][2] };
            w[0] = e1[1]*e2[2] - e1[2]*e2[1]; w[1] = e1[2]*e2[0] - e1[0]*e2[2]; w[2] = e1[0]*e2[1] - e1[1]*e2[0];
            ww = w[0]*w[0] + w[1]*w[1] + w[2]*w[2];
            vol = 0.5 * std::sqrt(ww);
            for( int r = 0; r < 3; ++r ) { for( int c = 0; c < 3; ++c ) { pr[r][c] = w[r] * w[c] / ww; } }
        }
        else
        {
            for( int k = 0; k < 3; ++k ) { w[k] = X[1][k] - X[0][k]; }
            ww = w[0]*w[0] + w[1]*w[1] + w[2]*w[2];
            vol = std::sqrt(ww);
            for( int r = 0; r < 3; ++r ) { for( int c = 0; c < 3; ++c ) { pr[r][c] = (r == c ? 1.0 : 0.0) - w[r] * w[c] / ww; } }
        }
        double mean[3] = {0,0,0};
        for( int l = 0; l < SIZE; ++l ) { for( int k = 0; k < 3; ++k ) { mean[k] += X[l][k] / SIZE; } }
        for( int j = 0; j < SIZE; ++j )
        {
            for( int c = 0; c < 3; ++c )
            {
                // dw = d w / d X_{j,c}
                double dw[3] = {0,0,0};
                if( n == 2 )
                {
                    // w = (X1-X0) x (X2-X0); d/dX1 = e_c x (X2-X0); d/dX2 = (X1-X0) x e_c; d/dX0 = -(sum).
                    const double e1[3] = { X[1][0]-X[0][0], X[1][1]-X[0][1], X[1][2]-X[0][2] };
                    const double e2[3] = { X[2][0]-X[0][0], X[2][1]-X[0][1], X[2][2]-X[0][2] };
                    double ec[3] = {0,0,0}; ec[c] = 1.0;
                    double a1[3] = { ec[1]*e2[2]-ec[2]*e2[1], ec[2]*e2[0]-ec[0]*e2[2], ec[0]*e2[1]-ec[1]*e2[0] };
                    double a2[3] = { e1[1]*ec[2]-e1[2]*ec[1], e1[2]*ec[0]-e1[0]*ec[2], e1[0]*ec[1]-e1[1]*ec[0] };
                    for( int k = 0; k < 3; ++k )
                    {
                        if( j == 1 ) { dw[k] = a1[k]; }
                        else if( j == 2 ) { dw[k] = a2[k]; }
                        else { dw[k] = -a1[k] - a2[k]; }
                    }
                }
                else
                {
                    dw[c] = (j == 1) ? 1.0 : -1.0;
                }
                const double wdw = w[0]*dw[0] + w[1]*dw[1] + w[2]*dw[2];
                const double dvol = (n == 2) ? 0.5 * wdw / std::sqrt(ww) : wdw / std::sqrt(ww);
                // d(w w^T / ww) = (dw w^T + w dw^T)/ww - 2 w w^T (w.dw)/ww^2
                double dpr[3][3];
                for( int r = 0; r < 3; ++r )
                {
                    for( int s = 0; s < 3; ++s )
                    {
                        const double d = (dw[r] * w[s] + w[r] * dw[s]) / ww - 2.0 * w[r] * w[s] * wdw / (ww * ww);
                        dpr[r][s] = (n == 2) ? d : -d;
                    }
                }
                double acc = (Nt[0] + Ft[0]) * dvol;
                for( int l = 0; l < SIZE; ++l )
                {
                    for( int k = 0; k < 3; ++k )
                    {
                        acc += Nt[1 + 3 * l + k] * ( dvol * X[l][k] + ((l == j && k == c) ? vol : 0.0) );
                    }
                }
                for( int k = 0; k < 3; ++k )
                {
                    acc += Ft[1 + k] * ( dvol * mean[k] + ((k == c) ? vol / SIZE : 0.0) );
                }
                for( int k = 0; k < 6; ++k )
                {
                    const double m = dvol * pr[pi[k]][pj[k]] + vol * dpr[pi[k]][pj[k]];
                    acc += (Nt[1 + 3 * SIZE + k] + Ft[4 + k]) * m;
                }
                B[(LInt)3 * SIZE * t + 3 * j + c] = charge * acc;
            }
        }
    }
}

// Y = alpha * A^T B + beta * Y, A = binary T(n+1) x V corner-to-vertex matrix (DerivativeAssembler).
// Summation per vertex in increasing corner index.
void orc_assemble( int n, LInt V, LInt T, const Int * S, const double * B, double alpha, double beta, double * Y, LInt ldY )
{
    const int SIZE = n + 1;
    std::vector<double> acc( (size_t)(3 * V), 0.0 );
    std::vector<char> touched( (size_t)V, 0 );
    for( LInt r = 0; r < T * SIZE; ++r )
    {
        const Int v = S[r];
        for( int k = 0; k < 3; ++k ) { acc[3 * (size_t)v + k] += B[3 * r + k]; }
    }
    for( LInt v = 0; v < V; ++v )
    {
        for( int k = 0; k < 3; ++k )
        {
            Y[ldY * v + k] = alpha * acc[3 * v + k] + (beta == 0.0 ? 0.0 : beta * Y[ldY * v + k]);
        }
    }
}

// Pre (MixedOrder) -> NF/FF block products -> Post, for 3 right-hand sides.
// pre_cols T x (n+1), pre_vals T x 4 x (n+1) (tree order; row 0 averaging, rows 1..3 derivative),
// nf_vals/ff_vals full-pattern (a0,a1), nf_diag T x 2, ff_diag C x 2.
// threads = 1 runs in order (the checker); more threads split the pre-operator and the block products by
// rows (cost-balanced) and the post-operator scatter into per-thread vertex buffers added in thread order.
void orc_multiply_metric_mt( int n, LInt V, LInt T, LInt C,
    const Int * C_left, const Int * C_right, const Int * C_begin, const Int * C_end,
    const Int * pre_cols, const double * pre_vals,
    const LInt * nrp, const Int * nci, const double * nf_vals, const double * nf_diag,
    const LInt * frp, const Int * fci, const double * ff_vals, const double * ff_diag,
    double alpha, const double * X, LInt ldX, double beta, double * Y, LInt ldY, int nf_flag, int ff_flag, int threads )
{
    const int SIZE = n + 1;
    threads = std::max( 1, threads );
    std::vector<double> P_in( (size_t)(12 * T), 0.0 ), P_out( (size_t)(12 * T), 0.0 );
    std::vector<double> C_in( (size_t)(12 * C), 0.0 ), C_out( (size_t)(12 * C), 0.0 );
    const std::vector<LInt> tcut = CostPartition( T, threads, []( LInt ) { return 0; } );
    ParallelRanges( tcut, [&]( int, LInt i0, LInt i1 ) { for( LInt i = i0; i < i1; ++i )
    {
        for( int r = 0; r < 4; ++r )
        {
            for( int k = 0; k < 3; ++k )
            {
                double s = 0;
                for( int l = 0; l < SIZE; ++l ) { s += pre_vals[4 * SIZE * i + SIZE * r + l] * X[ldX * pre_cols[SIZE * i + l] + k]; }
                P_in[12 * i + 3 * r + k] = s;
            }
        }
    } } );
    // PrimitivesToClusters + PercolateUp (children before parents: reverse depth-first order).
    for( LInt c = C; c-- > 0; )
    {
        if( C_left[c] < 0 )
        {
            for( int h = 0; h < 12; ++h ) { double s = 0; for( Int i = C_begin[c]; i < C_end[c]; ++i ) { s += P_in[12 * (size_t)i + h]; } C_in[12 * c + h] = s; }
        }
        else
        {
            for( int h = 0; h < 12; ++h ) { C_in[12 * c + h] = C_in[12 * C_left[c] + h] + C_in[12 * C_right[c] + h]; }
        }
    }
    auto apply = [&]( const LInt rows, const LInt * rp, const Int * ci, const double * vals, const double * diag,
                      const std::vector<double> & xin, std::vector<double> & yout )
    {
        ParallelRanges( CostPartition( rows, threads, [&]( LInt i ) { return rp[i + 1] - rp[i]; } ), [&]( int, LInt i0, LInt i1 ) {
        for( LInt i = i0; i < i1; ++i )
        {
            double y[12] = {0};
            for( LInt k = rp[i]; k < rp[i + 1]; ++k )
            {
                const double * x = &xin[12 * (size_t)ci[k]];
                const double a0 = vals[2 * k], a1 = vals[2 * k + 1];
                for( int h = 0; h < 3; ++h ) { y[h] += a0 * x[h]; }
                for( int h = 3; h < 12; ++h ) { y[h] += a1 * x[h]; }
            }
            const double * x = &xin[12 * (size_t)i];
            for( int h = 0; h < 3; ++h ) { y[h] += diag[2 * i] * x[h]; }
            for( int h = 3; h < 12; ++h ) { y[h] += diag[2 * i + 1] * x[h]; }
            for( int h = 0; h < 12; ++h ) { yout[12 * (size_t)i + h] = y[h]; }
        } } );
    };
    if( nf_flag && nrp[T] > 0 ) { apply( T, nrp, nci, nf_vals, nf_diag, P_in, P_out ); }
    if( ff_flag && frp[C] > 0 ) { apply( C, frp, fci, ff_vals, ff_diag, C_in, C_out ); }
    // Post: PercolateDown, ClustersToPrimitives(add), mi_post.
    for( LInt c = 0; c < C; ++c )
    {
        const Int L = C_left[c];
        if( L >= 0 )
        {
            const Int R = C_right[c];
            for( int h = 0; h < 12; ++h ) { C_out[12 * L + h] = C_out[12 * c + h] + C_out[12 * L + h]; }
            for( int h = 0; h < 12; ++h ) { C_out[12 * R + h] = C_out[12 * c + h] + C_out[12 * R + h]; }
        }
    }
    for( LInt c = 0; c < C; ++c )
    {
        if( C_left[c] < 0 )
        {
            for( Int i = C_begin[c]; i < C_end[c]; ++i ) { for( int h = 0; h < 12; ++h ) { P_out[12 * (size_t)i + h] += C_out[12 * c + h]; } }
        }
    }
    std::vector<std::vector<double>> accs( (size_t)threads );
    ParallelRanges( tcut, [&]( int t, LInt i0, LInt i1 )
    {
        std::vector<double> & acc = accs[(size_t)t];
        acc.assign( (size_t)(3 * V), 0.0 );
        for( LInt i = i0; i < i1; ++i )
        {
            for( int l = 0; l < SIZE; ++l )
            {
                const Int v = pre_cols[SIZE * i + l];
                for( int r = 0; r < 4; ++r )
                {
                    const double w = pre_vals[4 * SIZE * i + SIZE * r + l];
                    for( int k = 0; k < 3; ++k ) { acc[3 * (size_t)v + k] += w * P_out[12 * i + 3 * r + k]; }
                }
            }
        }
    });
    ReduceBuffers( accs[0].data(), accs, 3 * V, threads );
    const std::vector<double> & acc = accs[0];
    for( LInt v = 0; v < V; ++v )
    {
        for( int k = 0; k < 3; ++k ) { Y[ldY * v + k] = alpha * acc[3 * v + k] + (beta == 0.0 ? 0.0 : beta * Y[ldY * v + k]); }
    }
}

void orc_multiply_metric( int n, LInt V, LInt T, LInt C,
    const Int * C_left, const Int * C_right, const Int * C_begin, const Int * C_end,
    const Int * pre_cols, const double * pre_vals,
    const LInt * nrp, const Int * nci, const double * nf_vals, const double * nf_diag,
    const LInt * frp, const Int * fci, const double * ff_vals, const double * ff_diag,
    double alpha, const double * X, LInt ldX, double beta, double * Y, LInt ldY, int nf_flag, int ff_flag )
{
    orc_multiply_metric_mt( n, V, T, C, C_left, C_right, C_begin, C_end, pre_cols, pre_vals, nrp, nci, nf_vals, nf_diag,
                            frp, fci, ff_vals, ff_diag, alpha, X, ldX, beta, Y, ldY, nf_flag, ff_flag, 1 );
}

// Brute-force energy: sum over all unordered pairs i < j of a_i a_j E_ij from far rows
// (area, barycenter, projector), no hierarchy.
double orc_allpairs_energy( int n, LInt T, const double * P_far, double q, double p )
{
    const Exponents ex( q, p, n );
    Neumaier en;
    double DX[16], DY[16];
    for( LInt i = 0; i < T; ++i )
    {
        const double * X = P_far + 10 * i;
        for( LInt j = i + 1; j < T; ++j )
        {
            const double * Y = P_far + 10 * j;
            en.add( PairCompute( ex, n, true, false, false, X[0], X + 1, X + 4, Y[0], Y + 1, Y + 4, DX, DY, 1, 1, 1.0, 1.0 ).result );
        }
    }
    return en.value();
}

// Brute-force dense metric values for all ordered pairs i != j (T x T x 2), far rows as input.
void orc_allpairs_metric( int n, LInt T, const double * P_far, double q, double p, double * vals )
{
    const Exponents ex( q, p, n );
    double DX[16], DY[16];
    for( LInt i = 0; i < T; ++i )
    {
        const double * X = P_far + 10 * i;
        for( LInt j = 0; j < T; ++j )
        {
            if( i == j ) { vals[2 * (T * i + j)] = 0; vals[2 * (T * i + j) + 1] = 0; continue; }
            const double * Y = P_far + 10 * j;
            const PairOut o = PairCompute( ex, n, false, false, true, X[0], X + 1, X + 4, Y[0], Y + 1, Y + 4, DX, DY, 1, 1, 1.0, 1.0 );
            vals[2 * (T * i + j)] = o.a0; vals[2 * (T * i + j) + 1] = o.a1;
        }
    }
}

// PseudoLaplacian_Traversor::Compute (LowOrder) for a symmetric BCT: on the upper part of each
// pattern, val = -Power(r2, s_exp) with s_exp = -n/2 - s (PseudoLaplacian_Kernel_NF.hpp:124-146,
// _FF.hpp:101-146), row diagonal -= (b/a) val, column diagonal += -(a/b) val; the lower part is
// filled from the upper part. Near rows use the barycenter of the hull (FMM_Kernel_NF::loadS),
// far rows the cluster center. nf_vals N_nf, ff_vals N_ff, nf_diag T, ff_diag C.
void orc_pl_compute_mt( int n, LInt T, LInt C, const double * P_near, const double * C_far,
    const LInt * nrp, const Int * nci, const LInt * ndiag,
    const LInt * frp, const Int * fci, const LInt * fdiag,
    double s, double * nf_vals, double * ff_vals, double * nf_diag, double * ff_diag, int threads )
{
    threads = std::max( 1, threads );
    const int SIZE = n + 1;
    const int ND = 1 + 3 * SIZE + 6;
    const double scale = 1.0 / SIZE;
    const double s_exp = -0.5 * n - s;
    auto power = [&]( double r2 )
    {
        const double k = std::nearbyint( s_exp );
        return ( k == s_exp ) ? IntPow( r2, (LInt)k ) : std::pow( r2, s_exp );
    };
    auto run = [&]( LInt rows, const LInt * rp, const Int * ci, const LInt * dg, double * vals, double * diag,
                    auto && load )
    {
        std::memset( diag, 0, sizeof(double) * (size_t)rows );
        std::vector<std::vector<double>> gbuf( (size_t)threads );
        ParallelRanges( CostPartition( rows, threads, [&]( LInt i ) { return rp[i + 1] - dg[i]; } ), [&]( int t, LInt i0, LInt i1 )
        {
            double * D = diag;
            if( t > 0 ) { gbuf[(size_t)t].assign( (size_t)rows, 0.0 ); D = gbuf[(size_t)t].data(); }
            for( LInt i = i0; i < i1; ++i )
            {
                double a, x[3];
                load( i, a, x );
                double ii_block = 0;
                for( LInt k = dg[i]; k < rp[i + 1]; ++k )
                {
                    const Int j = ci[k];
                    double b, y[3];
                    load( j, b, y );
                    double r2 = 0;
                    for( int l = 0; l < 3; ++l ) { const double v = y[l] - x[l]; r2 += v * v; }
                    const double val = -power( r2 );
                    vals[k] = val;
                    ii_block -= (b / a) * val;
                    D[j] += -(a / b) * val;
                }
                D[i] += ii_block;
            }
        });
        ReduceBuffers( diag, gbuf, rows, threads );
        ParallelRanges( CostPartition( rows, threads, [&]( LInt i ) { return dg[i] - rp[i]; } ), [&]( int, LInt i0, LInt i1 )
        {
            for( LInt i = i0; i < i1; ++i )
            {
                for( LInt k = rp[i]; k < dg[i]; ++k )
                {
                    const Int j = ci[k];
                    const Int * f = std::lower_bound( ci + dg[j], ci + rp[j + 1], (Int)i );
                    vals[k] = vals[f - ci];
                }
            }
        });
    };
    run( T, nrp, nci, ndiag, nf_vals, nf_diag, [&]( LInt i, double & a, double * x )
    {
        const double * X = P_near + ND * i;
        a = X[0];
        for( int k = 0; k < 3; ++k ) { x[k] = scale * X[1 + k]; for( int kk = 1; kk < SIZE; ++kk ) { x[k] += scale * X[1 + 3 * kk + k]; } }
    });
    run( C, frp, fci, fdiag, ff_vals, ff_diag, [&]( LInt i, double & a, double * x )
    {
        const double * X = C_far + 10 * i;
        a = X[0];
        for( int k = 0; k < 3; ++k ) { x[k] = X[1 + k]; }
    });
}

void orc_pl_compute( int n, LInt T, LInt C, const double * P_near, const double * C_far,
    const LInt * nrp, const Int * nci, const LInt * ndiag,
    const LInt * frp, const Int * fci, const LInt * fdiag,
    double s, double * nf_vals, double * ff_vals, double * nf_diag, double * ff_diag )
{
    orc_pl_compute_mt( n, T, C, P_near, C_far, nrp, nci, ndiag, frp, fci, fdiag, s, nf_vals, ff_vals, nf_diag, ff_diag, 1 );
}

// MultiplyMetric of the pseudo-Laplacian: LowOrder Pre (row 0 of the pre-operator: a/(n+1) per
// vertex, ClusterTree.hpp:458-500) -> scalar NF/FF block products plus diagonals -> Post
// (PercolateDown, ClustersToPrimitives, lo_post = lo_pre^T), for 3 right-hand sides.
void orc_pl_multiply_mt( int n, LInt V, LInt T, LInt C,
    const Int * C_left, const Int * C_right, const Int * C_begin, const Int * C_end,
    const Int * pre_cols, const double * pre_vals,
    const LInt * nrp, const Int * nci, const double * nf_vals, const double * nf_diag,
    const LInt * frp, const Int * fci, const double * ff_vals, const double * ff_diag,
    double alpha, const double * X, LInt ldX, double beta, double * Y, LInt ldY, int nf_flag, int ff_flag, int threads )
{
    const int SIZE = n + 1;
    threads = std::max( 1, threads );
    std::vector<double> P_in( (size_t)(3 * T), 0.0 ), P_out( (size_t)(3 * T), 0.0 );
    std::vector<double> C_in( (size_t)(3 * C), 0.0 ), C_out( (size_t)(3 * C), 0.0 );
    const std::vector<LInt> tcut = CostPartition( T, threads, []( LInt ) { return 0; } );
    ParallelRanges( tcut, [&]( int, LInt i0, LInt i1 )
    {
        for( LInt i = i0; i < i1; ++i )
        {
            for( int k = 0; k < 3; ++k )
            {
                double s = 0;
                for( int l = 0; l < SIZE; ++l ) { s += pre_vals[4 * SIZE * i + l] * X[ldX * pre_cols[SIZE * i + l] + k]; }
                P_in[3 * i + k] = s;
            }
        }
    });
    for( LInt c = C; c-- > 0; )
    {
        if( C_left[c] < 0 )
        {
            for( int h = 0; h < 3; ++h ) { double s = 0; for( Int i = C_begin[c]; i < C_end[c]; ++i ) { s += P_in[3 * (size_t)i + h]; } C_in[3 * c + h] = s; }
        }
        else
        {
            for( int h = 0; h < 3; ++h ) { C_in[3 * c + h] = C_in[3 * C_left[c] + h] + C_in[3 * C_right[c] + h]; }
        }
    }
    auto apply = [&]( const LInt rows, const LInt * rp, const Int * ci, const double * vals, const double * diag,
                      const std::vector<double> & xin, std::vector<double> & yout )
    {
        ParallelRanges( CostPartition( rows, threads, [&]( LInt i ) { return rp[i + 1] - rp[i]; } ), [&]( int, LInt i0, LInt i1 )
        {
            for( LInt i = i0; i < i1; ++i )
            {
                double y[3] = {0, 0, 0};
                for( LInt k = rp[i]; k < rp[i + 1]; ++k )
                {
                    const double * x = &xin[3 * (size_t)ci[k]];
                    for( int h = 0; h < 3; ++h ) { y[h] += vals[k] * x[h]; }
                }
                for( int h = 0; h < 3; ++h ) { yout[3 * (size_t)i + h] = y[h] + diag[i] * xin[3 * (size_t)i + h]; }
            }
        });
    };
    if( nf_flag && nrp[T] > 0 ) { apply( T, nrp, nci, nf_vals, nf_diag, P_in, P_out ); }
    if( ff_flag && frp[C] > 0 ) { apply( C, frp, fci, ff_vals, ff_diag, C_in, C_out ); }
    for( LInt c = 0; c < C; ++c )
    {
        const Int L = C_left[c];
        if( L >= 0 )
        {
            const Int R = C_right[c];
            for( int h = 0; h < 3; ++h ) { C_out[3 * L + h] = C_out[3 * c + h] + C_out[3 * L + h]; }
            for( int h = 0; h < 3; ++h ) { C_out[3 * R + h] = C_out[3 * c + h] + C_out[3 * R + h]; }
        }
    }
    for( LInt c = 0; c < C; ++c )
    {
        if( C_left[c] < 0 )
        {
            for( Int i = C_begin[c]; i < C_end[c]; ++i ) { for( int h = 0; h < 3; ++h ) { P_out[3 * (size_t)i + h] += C_out[3 * c + h]; } }
        }
    }
    std::vector<std::vector<double>> accs( (size_t)threads );
    ParallelRanges( tcut, [&]( int t, LInt i0, LInt i1 )
    {
        std::vector<double> & acc = accs[(size_t)t];
        acc.assign( (size_t)(3 * V), 0.0 );
        for( LInt i = i0; i < i1; ++i )
        {
            for( int l = 0; l < SIZE; ++l )
            {
                const Int v = pre_cols[SIZE * i + l];
                const double w = pre_vals[4 * SIZE * i + l];
                for( int k = 0; k < 3; ++k ) { acc[3 * (size_t)v + k] += w * P_out[3 * i + k]; }
            }
        }
    });
    ReduceBuffers( accs[0].data(), accs, 3 * V, threads );
    const std::vector<double> & acc = accs[0];
    for( LInt v = 0; v < V; ++v )
    {
        for( int k = 0; k < 3; ++k ) { Y[ldY * v + k] = alpha * acc[3 * v + k] + (beta == 0.0 ? 0.0 : beta * Y[ldY * v + k]); }
    }
}

void orc_pl_multiply( int n, LInt V, LInt T, LInt C,
    const Int * C_left, const Int * C_right, const Int * C_begin, const Int * C_end,
    const Int * pre_cols, const double * pre_vals,
    const LInt * nrp, const Int * nci, const double * nf_vals, const double * nf_diag,
    const LInt * frp, const Int * fci, const double * ff_vals, const double * ff_diag,
    double alpha, const double * X, LInt ldX, double beta, double * Y, LInt ldY, int nf_flag, int ff_flag )
{
    orc_pl_multiply_mt( n, V, T, C, C_left, C_right, C_begin, C_end, pre_cols, pre_vals, nrp, nci, nf_vals, nf_diag,
                        frp, fci, ff_vals, ff_diag, alpha, X, ldX, beta, Y, ldY, nf_flag, ff_flag, 1 );
}

} // extern "C"

// ------------------------------------------------------------------------------------------
// Very-near field (VF): adaptive subdivision quadrature for primitive pairs that are neither
// adjacent nor separated by the near-field MAC.
//   FMM_Kernel_VF_Common::Compute        src/FMM/FMM_Kernel_VF_Common.hpp:3-116
//   FMM_Kernel_VF::loadS/loadT/writeS/writeT src/FMM/FMM_Kernel_VF.hpp:243-305
//   SimplexHierarchy RequireSimplex/RequireSubsimplex src/SimplexHierarchy/SimplexHierarchy_Details.hpp,
//                    ToChild/ToParent src/SimplexHierarchy/SimplexHierarchy_1D.hpp, _2D.hpp
//   TP0_Kernel_VF::compute               src/Energies/TangentPointEnergy0/TP0_Kernel_VF.hpp:158-338
//   PseudoLaplacian_Kernel_VF::compute   src/Metrics/PseudoLaplacian/PseudoLaplacian_Kernel_VF.hpp:135-176
// The GJK multipole acceptance criterion (src/GJK.hpp:1162-1172) is restated with the exact
// simplex distance: separated <=> max(r_S^2, r_T^2) < theta^2 dist(S,T)^2, with r^2 of a sub-simplex
// = scale^2 * r^2 of its primitive (max squared vertex distance from the barycenter).
namespace
{
    struct V3o { double x, y, z; };
    inline V3o sub3( V3o a, V3o b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
    inline V3o add3( V3o a, V3o b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
    inline V3o mul3( double s, V3o a ) { return { s * a.x, s * a.y, s * a.z }; }
    inline double dot3( V3o a, V3o b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    inline double clamp01o( double t ) { return t < 0 ? 0 : (t > 1 ? 1 : t); }

    // Closest point on triangle abc to p by Voronoi regions of the vertices, edges and face.
    V3o ClosestOnTriangle( V3o p, V3o a, V3o b, V3o c )
    {
        const V3o ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
        const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
        if( d1 <= 0 && d2 <= 0 ) { return a; }
        const V3o bp = sub3(p, b);
        const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
        if( d3 >= 0 && d4 <= d3 ) { return b; }
        const double vc = d1 * d4 - d3 * d2;
        if( vc <= 0 && d1 >= 0 && d3 <= 0 ) { return add3( a, mul3( d1 / (d1 - d3), ab ) ); }
        const V3o cp = sub3(p, c);
        const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
        if( d6 >= 0 && d5 <= d6 ) { return c; }
        const double vb = d5 * d2 - d1 * d6;
        if( vb <= 0 && d2 >= 0 && d6 <= 0 ) { return add3( a, mul3( d2 / (d2 - d6), ac ) ); }
        const double va = d3 * d6 - d5 * d4;
        if( va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0 ) { return add3( b, mul3( (d4 - d3) / ((d4 - d3) + (d5 - d6)), sub3(c, b) ) ); }
        const double den = 1.0 / (va + vb + vc);
        return add3( a, add3( mul3( vb * den, ab ), mul3( vc * den, ac ) ) );
    }

    // Squared distance of the segments [p1,q1], [p2,q2] (clamped parametric closest points).
    double SegmentsDist2( V3o p1, V3o q1, V3o p2, V3o q2 )
    {
        const V3o d1 = sub3(q1, p1), d2 = sub3(q2, p2), r = sub3(p1, p2);
        const double a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
        double s = 0, t = 0;
        if( a <= 1e-300 && e <= 1e-300 ) { return dot3(r, r); }
        if( a <= 1e-300 ) { t = clamp01o( f / e ); }
        else
        {
            const double c = dot3(d1, r);
            if( e <= 1e-300 ) { s = clamp01o( -c / a ); }
            else
            {
                const double b = dot3(d1, d2), den = a * e - b * b;
                s = (den != 0) ? clamp01o( (b * f - c * e) / den ) : 0.0;
                t = (b * s + f) / e;
                if( t < 0 ) { t = 0; s = clamp01o( -c / a ); }
                else if( t > 1 ) { t = 1; s = clamp01o( (b - c) / a ); }
            }
        }
        const V3o w = sub3( add3( p1, mul3( s, d1 ) ), add3( p2, mul3( t, d2 ) ) );
        return dot3(w, w);
    }

    // Squared distance of two disjoint simplices with m = 2 (segments) or 3 (triangles) points each.
    double SimplexDist2( const V3o * P, const V3o * Q, int m )
    {
        double d2 = 1e300;
        if( m == 2 ) { return SegmentsDist2( P[0], P[1], Q[0], Q[1] ); }
        for( int i = 0; i < 3; ++i )
        {
            V3o c = ClosestOnTriangle( P[i], Q[0], Q[1], Q[2] ); d2 = std::min( d2, dot3( sub3(P[i], c), sub3(P[i], c) ) );
            c = ClosestOnTriangle( Q[i], P[0], P[1], P[2] );     d2 = std::min( d2, dot3( sub3(Q[i], c), sub3(Q[i], c) ) );
        }
        for( int i = 0; i < 3; ++i ) { for( int j = 0; j < 3; ++j ) { d2 = std::min( d2, SegmentsDist2( P[i], P[(i+1)%3], Q[j], Q[(j+1)%3] ) ); } }
        return d2;
    }

    // Sub-simplex of one primitive: barycentric corners and center, level, child ids, path code.
    struct Hier
    {
        int n = 2;
        double corners[3][3], center[3];
        int level = 0, child = 0, former = -1;
        uint64_t column = 0;
        double scale = 1, weight = 1;

        int Children() const { return n == 2 ? 4 : 2; }
        void Reset()
        {
            const int S = n + 1;
            for( int i = 0; i < S; ++i ) { center[i] = 1.0 / S; for( int j = 0; j < S; ++j ) { corners[i][j] = (i == j) ? 1.0 : 0.0; } }
            level = 0; child = 0; former = -1; column = 0; scale = 1; weight = 1;
        }
        // Child k < n+1 is the half-size copy at corner k; child 3 of a triangle the inverted middle.
        void ToChild( int k )
        {
            const int S = n + 1;
            former = child; child = k; column = (uint64_t)Children() * column + (uint64_t)k; ++level;
            scale *= 0.5; weight *= (n == 2 ? 0.25 : 0.5);
            if( k < S )
            {
                for( int i = 0; i < S; ++i )
                {
                    if( i != k ) { for( int j = 0; j < S; ++j ) { corners[i][j] = corners[k][j] + 0.5 * (corners[i][j] - corners[k][j]); } }
                    else         { for( int j = 0; j < S; ++j ) { center[j] = corners[k][j] + 0.5 * (center[j] - corners[k][j]); } }
                }
            }
            else
            {
                for( int i = 0; i < S; ++i ) { for( int j = 0; j < S; ++j ) { corners[i][j] = center[j] - 0.5 * (corners[i][j] - center[j]); } }
            }
        }
        void ToParent()
        {
            const int S = n + 1;
            const int k = child;
            column = (column - (uint64_t)child) / (uint64_t)Children();
            former = child; child = (int)(column % (uint64_t)Children()); --level;
            scale *= 2.0; weight *= (n == 2 ? 4.0 : 2.0);
            if( k < S )
            {
                for( int i = 0; i < S; ++i )
                {
                    if( i != k ) { for( int j = 0; j < S; ++j ) { corners[i][j] = corners[k][j] + 2.0 * (corners[i][j] - corners[k][j]); } }
                    else         { for( int j = 0; j < S; ++j ) { center[j] = corners[k][j] + 2.0 * (center[j] - corners[k][j]); } }
                }
            }
            else
            {
                for( int i = 0; i < S; ++i ) { for( int j = 0; j < S; ++j ) { corners[i][j] = center[j] - 2.0 * (corners[i][j] - center[j]); } }
            }
        }
        // Vertices of the sub-simplex in space from the primitive's hull (SIZE x 3).
        void Vertices( const double * hull, V3o * out ) const
        {
            const int S = n + 1;
            for( int i = 0; i < S; ++i )
            {
                double c[3];
                for( int k = 0; k < 3; ++k ) { c[k] = corners[i][0] * hull[k]; for( int j = 1; j < S; ++j ) { c[k] += corners[i][j] * hull[3 * j + k]; } }
                out[i] = { c[0], c[1], c[2] };
            }
        }
    };

    // Squared radius of a primitive as the reference serializes it (Polytope::FromCoordinates,
    // src/Primitives/Polytope.hpp:69-111): max squared vertex distance from the vertex average.
    double PrimitiveSquaredRadius( const double * hull, int S )
    {
        double c[3];
        for( int k = 0; k < 3; ++k ) { c[k] = hull[k]; for( int j = 1; j < S; ++j ) { c[k] += hull[3 * j + k]; } c[k] *= 1.0 / S; }
        double r2 = 0;
        for( int j = 0; j < S; ++j )
        {
            double d = hull[3 * j] - c[0];
            double sq = d * d;
            for( int k = 1; k < 3; ++k ) { d = hull[3 * j + k] - c[k]; sq += d * d; }
            r2 = std::max( r2, sq );
        }
        return r2;
    }

    // Depth-first traversal of FMM_Kernel_VF_Common::Compute for one primitive pair: calls
    // eval(S, T) for every admissible sub-pair or every sub-pair at level max_refinement.
    template<typename Eval>
    void VFTraverse( int n, const double * hullS, const double * hullT, double r2S, double r2T, double theta2, int max_refinement,
                     Hier & S, Hier & T, double * stats, Eval && eval )
    {
        S.n = n; T.n = n; S.Reset(); T.Reset();
        const int m = n + 1;
        bool from_above = true, go_on = true;
        S.ToChild(0); T.ToChild(0);
        while( go_on )
        {
            if( from_above )
            {
                bool leaf = S.level >= max_refinement;
                if( !leaf )
                {
                    V3o PS[3], PT[3];
                    S.Vertices( hullS, PS ); T.Vertices( hullT, PT );
                    const double rr = std::max( S.scale * S.scale * r2S, T.scale * T.scale * r2T );
                    leaf = rr < theta2 * SimplexDist2( PS, PT, m );
                }
                if( leaf )
                {
                    stats[0] += 1.0; stats[1] = std::max( stats[1], (double)std::min( S.level, max_refinement ) );
                    eval( S, T );
                    S.ToParent(); T.ToParent(); from_above = false;
                }
                else { S.ToChild(0); T.ToChild(0); from_above = true; }
            }
            else
            {
                const int sk = S.former, tk = T.former;
                if( tk + 1 < T.Children() ) { S.ToChild( sk ); T.ToChild( tk + 1 ); from_above = true; }
                else if( sk + 1 < S.Children() ) { S.ToChild( sk + 1 ); T.ToChild( 0 ); from_above = true; }
                else if( S.level == 0 ) { go_on = false; }
                else { S.ToParent(); T.ToParent(); from_above = false; }
            }
        }
    }
}

extern "C"
{

// Squared distance of two disjoint simplices with m = 2 (segments) or 3 (triangles) points, row-major.
double orc_simplex_dist2( int m, const double * P, const double * Q )
{
    V3o p[3], q[3];
    for( int i = 0; i < m; ++i ) { p[i] = { P[3*i], P[3*i+1], P[3*i+2] }; q[i] = { Q[3*i], Q[3*i+1], Q[3*i+2] }; }
    return SimplexDist2( p, q, m );
}

// TP0_Traversor::VF_Compute for a symmetric BCT: upper part of the full symmetric VeryNear pattern
// (rp, ci, diag as for the near field). P_near: T x ND (tree order). flags: 1 energy, 2 differential,
// 4 metric, 8 multiply each energy term by the quadrature weight w (not what the reference does:
// TP0_Kernel_VF.hpp:273 sums a E b without w; the weighted form is the quadrature whose derivative
// the differential is, used to pin the differential). Outputs added into P_D_near (T x ND) and
// vf_diag (T x 2, the first two entries of the reference's 16-value diagonal blocks); vf_vals
// (full pattern x 2) receives (a0, a1) per entry, the lower part copied from the upper part.
// stats[0] += evaluations, stats[1] = max(stats[1], deepest level). Returns the energy.
double orc_vf_compute( int n, LInt T, const double * P_near, const LInt * rp, const Int * ci, const LInt * diag,
    double theta, int max_refinement, double q, double p, int flags,
    double * P_D_near, double * vf_vals, double * vf_diag, double * stats )
{
    const int SIZE = n + 1;
    const int ND = 1 + 3 * SIZE + 6;
    const bool E = flags & 1, D = flags & 2, M = flags & 4, W = flags & 8;
    const Exponents ex( q, p, n );
    const double theta2 = theta * theta;
    Neumaier en;
    std::vector<double> r2( (size_t)T );
    for( LInt i = 0; i < T; ++i ) { r2[i] = PrimitiveSquaredRadius( P_near + ND * i + 1, SIZE ); }
    Hier S, Tt;
    for( LInt i = 0; i < T; ++i )
    {
        if( rp[i + 1] <= diag[i] ) { continue; }
        const double * X = P_near + ND * i;
        const double a = X[0];
        const double * xb = X + 1, * P = X + 1 + 3 * SIZE;
        double DX[16] = {0}, ii0 = 0, ii1 = 0;
        for( LInt k = diag[i]; k < rp[i + 1]; ++k )
        {
            const Int j = ci[k];
            const double * Y = P_near + ND * j;
            const double b = Y[0];
            const double * yb = Y + 1, * Q = Y + 1 + 3 * SIZE;
            double DY[16] = {0}, ij0 = 0, ij1 = 0, jj0 = 0, jj1 = 0;
            double sum = 0;
            VFTraverse( n, xb, yb, r2[i], r2[j], theta2, max_refinement, S, Tt, stats, [&]( const Hier & hs, const Hier & ht )
            {
                const double w = hs.weight * ht.weight;
                double x[3], y[3], v[3], Pv[3] = {0,0,0}, Qv[3] = {0,0,0}, Vp[6] = {0,0,0,0,0,0};
                double rr2 = 0, phi = 0, psi = 0;
                for( int c = 0; c < 3; ++c )
                {
                    x[c] = hs.center[0] * xb[c]; y[c] = ht.center[0] * yb[c];
                    for( int h = 1; h < SIZE; ++h ) { x[c] += hs.center[h] * xb[3 * h + c]; }
                    for( int h = 1; h < SIZE; ++h ) { y[c] += ht.center[h] * yb[3 * h + c]; }
                    v[c] = y[c] - x[c]; rr2 += v[c] * v[c];
                }
                for( int r = 0; r < 3; ++r )
                {
                    for( int c = 0; c < 3; ++c )
                    {
                        const int kk = lin_k3(r, c);
                        Pv[r] += P[kk] * v[c]; Qv[r] += Q[kk] * v[c];
                        if( c >= r ) { Vp[kk] = (1.0 + (double)(r != c)) * v[r] * v[c]; }
                    }
                    phi += v[r] * Pv[r]; psi += v[r] * Qv[r];
                }
                if( E || D )
                {
                    const double rmp2 = ex.PowP( rr2, -ex.p_half - 1, -ex.p_half_i - 1 );
                    const double rmp = rmp2 * rr2;
                    double phq2, psq2, phq, psq;
                    if( ex.q_flag >= 2 )
                    {
                        phq2 = ex.PowQ( std::abs(phi), ex.q_half - 1, ex.q_half_i - 1 );
                        psq2 = ex.PowQ( std::abs(psi), ex.q_half - 1, ex.q_half_i - 1 );
                        phq = phq2 * phi; psq = psq2 * psi;
                    }
                    else
                    {
                        phq2 = 1; psq2 = 1;
                        phq = ex.PowQ( std::abs(phi), ex.q_half, ex.q_half_i );
                        psq = ex.PowQ( std::abs(psi), ex.q_half, ex.q_half_i );
                    }
                    const double Num = phq + psq;
                    const double Ev = Num * rmp;
                    if( E ) { sum += W ? w * a * Ev * b : a * Ev * b; }
                    if( D )
                    {
                        const double factor = (ex.q_flag == 0) ? 0.0 : ex.q * rmp;
                        const double Kxy = (ex.q_flag == 0) ? 0.0 : factor * phq2;
                        const double Kyx = (ex.q_flag == 0) ? 0.0 : factor * psq2;
                        const double H = ex.p * rmp2 * Num;
                        const double wa = w * a, wb = w * b;
                        double dx = 0, dy = 0;
                        for( int c = 0; c < 3; ++c )
                        {
                            const double g = ((ex.q_flag == 0) ? 0.0 : Kxy * Pv[c] + Kyx * Qv[c]) - H * v[c];
                            dx += g * x[c]; dy += g * y[c];
                            for( int h = 0; h < SIZE; ++h ) { DX[1 + 3 * h + c] -= hs.center[h] * wb * g; }
                            for( int h = 0; h < SIZE; ++h ) { DY[1 + 3 * h + c] += ht.center[h] * wa * g; }
                        }
                        DX[0] += wb * ( Ev - 0.5 * factor * phq + dx );
                        DY[0] += wa * ( Ev - 0.5 * factor * psq - dy );
                        for( int kk = 0; kk < 6; ++kk )
                        {
                            DX[1 + 3 * SIZE + kk] += 0.5 * wb * Kxy * Vp[kk];
                            DY[1 + 3 * SIZE + kk] += 0.5 * wa * Kyx * Vp[kk];
                        }
                    }
                }
                if( M )
                {
                    const double r4 = rr2 * rr2;
                    const double a1 = w * -2.0 * std::pow( rr2, ex.s_exp );
                    const double a0 = w * 0.5 * (phi + psi) / r4 * a1;     // w enters twice, as in TP0_Kernel_VF.hpp:320
                    ij0 += a0; ij1 += a1;
                    ii0 -= (b / a) * a0; ii1 -= (b / a) * a1;
                    jj0 -= (a / b) * a0; jj1 -= (a / b) * a1;
                }
            });
            en.add( sum );
            if( D ) { for( int h = 0; h < ND; ++h ) { P_D_near[ND * j + h] += DY[h]; } }
            if( M )
            {
                vf_vals[2 * k] = ij0; vf_vals[2 * k + 1] = ij1;
                vf_diag[2 * j] += jj0; vf_diag[2 * j + 1] += jj1;
            }
        }
        if( D ) { for( int h = 0; h < ND; ++h ) { P_D_near[ND * i + h] += DX[h]; } }
        if( M ) { vf_diag[2 * i] += ii0; vf_diag[2 * i + 1] += ii1; }
    }
    if( M )
    {
        // lower entries from their transposed upper entries
        for( LInt i = 0; i < T; ++i )
        {
            for( LInt k = rp[i]; k < diag[i]; ++k )
            {
                const Int j = ci[k];
                const Int * f = std::lower_bound( ci + diag[j], ci + rp[j + 1], (Int)i );
                vf_vals[2 * k] = vf_vals[2 * (f - ci)]; vf_vals[2 * k + 1] = vf_vals[2 * (f - ci) + 1];
            }
        }
    }
    return E ? en.value() : 0.0;
}

// PseudoLaplacian_Traversor::VF_Compute (LowOrder): val = -w Power(r2, s_exp), s_exp = -n/2 - s,
// summed over the sub-pairs of the adaptive subdivision (max_refinement 20 in the reference,
// PseudoLaplacian_Traversor.hpp:187); row diagonal -= (b/a) val, column diagonal -= (a/b) val.
// vf_vals: full pattern (lower copied from upper), vf_diag T (added).
void orc_pl_vf( int n, LInt T, const double * P_near, const LInt * rp, const Int * ci, const LInt * diag,
                double theta, int max_refinement, double s, double * vf_vals, double * vf_diag, double * stats )
{
    const int SIZE = n + 1;
    const int ND = 1 + 3 * SIZE + 6;
    const double s_exp = -0.5 * n - s;
    const double theta2 = theta * theta;
    std::vector<double> r2( (size_t)T );
    for( LInt i = 0; i < T; ++i ) { r2[i] = PrimitiveSquaredRadius( P_near + ND * i + 1, SIZE ); }
    Hier S, Tt;
    for( LInt i = 0; i < T; ++i )
    {
        const double a = P_near[ND * i];
        double ii = 0;
        for( LInt k = diag[i]; k < rp[i + 1]; ++k )
        {
            const Int j = ci[k];
            const double b = P_near[ND * j];
            double ij = 0, jj = 0;
            const double * xb = P_near + ND * i + 1, * yb = P_near + ND * j + 1;
            VFTraverse( n, xb, yb, r2[i], r2[j], theta2, max_refinement, S, Tt, stats, [&]( const Hier & hs, const Hier & ht )
            {
                const double w = hs.weight * ht.weight;
                double rr2 = 0;
                for( int c = 0; c < 3; ++c )
                {
                    double x = hs.center[0] * xb[c], y = ht.center[0] * yb[c];
                    for( int h = 1; h < SIZE; ++h ) { x += hs.center[h] * xb[3 * h + c]; y += ht.center[h] * yb[3 * h + c]; }
                    rr2 += (y - x) * (y - x);
                }
                const double val = -w * std::pow( rr2, s_exp );
                ij += val; ii -= (b / a) * val; jj -= (a / b) * val;
            });
            vf_vals[k] = ij;
            vf_diag[j] += jj;
        }
        vf_diag[i] += ii;
    }
    for( LInt i = 0; i < T; ++i )
    {
        for( LInt k = rp[i]; k < diag[i]; ++k )
        {
            const Int j = ci[k];
            const Int * f = std::lower_bound( ci + diag[j], ci + rp[j + 1], (Int)i );
            vf_vals[k] = vf_vals[f - ci];
        }
    }
}

} // extern "C"

// ------------------------------------------------------------------------------------------
// MaximumSafeStepSize (the line-search bound after the metric solve):
//   SimplicialMesh::MaximumSafeStepSize           src/SimplicialMesh.hpp:441-468
//   CollisionTree::MaximumSafeStepSize           src/CollisionTree.hpp (minimum over the kernels)
//   MaximumSafeStepSize_Kernel::ComputeLeaf      src/CollisionTree/MaximumSafeStepSize_Kernel.hpp
//                                                (adjacent primitives are skipped)
//   CollisionFinder::FindMaximumSafeStepSize     src/CollisionFinder.hpp (bisection of [0, t_init])
//   MovingPolytope                               src/Primitives/MovingPolytope.hpp: the space-time convex
//                                                hull of the simplex at times a and b
//   GJK::IntersectingQ                           src/GJK.hpp:777-797, restated as an exact test: the two
//                                                hulls meet iff sum l_k P_k = sum m_k Q_k, sum l = sum m = 1,
//                                                l, m >= 0 is feasible (phase-one simplex, Bland's rule)
// Every pair starts its bisection at t_init = t_max (the reference starts later pairs at its running
// minimum, which changes results only within the tolerance).
namespace
{
    // Feasibility of { x >= 0 : A x = b } (R rows, N columns, b >= 0) by phase one of the simplex method.
    bool LPFeasible( int R, int N, const double * A, const double * b )
    {
        const int W = N + R + 1;
        std::vector<double> t( (size_t)( (R + 1) * W ), 0.0 );
        std::vector<int> basis( (size_t)R );
        for( int r = 0; r < R; ++r )
        {
            for( int c = 0; c < N; ++c ) { t[(size_t)(r * W + c)] = A[r * N + c]; }
            t[(size_t)(r * W + N + r)] = 1.0;
            t[(size_t)(r * W + W - 1)] = b[r];
            basis[(size_t)r] = N + r;
        }
        // objective row: minimize the sum of the artificials -> reduced costs
        double * obj = &t[(size_t)(R * W)];
        for( int r = 0; r < R; ++r ) { for( int c = 0; c < W; ++c ) { if( c < N || c == W - 1 ) { obj[c] -= t[(size_t)(r * W + c)]; } } }
        double scale = 1.0;
        for( int k = 0; k < R * N; ++k ) { scale = std::max( scale, std::abs( A[k] ) ); }
        const double eps = 1e-12 * scale;
        for( int it = 0; it < 500; ++it )
        {
            int enter = -1;
            for( int c = 0; c < N + R; ++c ) { if( obj[c] < -eps ) { enter = c; break; } }   // Bland: first improving column
            if( enter < 0 ) { break; }
            int leave = -1; double best = 0;
            for( int r = 0; r < R; ++r )
            {
                const double a = t[(size_t)(r * W + enter)];
                if( a > eps )
                {
                    const double ratio = t[(size_t)(r * W + W - 1)] / a;
                    if( leave < 0 || ratio < best - 1e-15 || ( ratio <= best + 1e-15 && basis[(size_t)r] < basis[(size_t)leave] ) ) { leave = r; best = ratio; }
                }
            }
            if( leave < 0 ) { break; }
            const double piv = t[(size_t)(leave * W + enter)];
            for( int c = 0; c < W; ++c ) { t[(size_t)(leave * W + c)] /= piv; }
            for( int r = 0; r <= R; ++r )
            {
                if( r == leave ) { continue; }
                const double f = t[(size_t)(r * W + enter)];
                if( f != 0.0 ) { for( int c = 0; c < W; ++c ) { t[(size_t)(r * W + c)] -= f * t[(size_t)(leave * W + c)]; } }
            }
            basis[(size_t)leave] = enter;
        }
        return -obj[W - 1] <= 1e-10;
    }

    // Do the space-time hulls of the two moving simplices (m points each) over [a, b] meet?
    bool MovingHullsMeet( int m, const double * p, const double * u, const double * q, const double * v, double a, double b, double Tscale )
    {
        // columns: l (2m) then mu (2m); rows: sum l = 1, sum mu = 1, 4 coordinates
        const int N = 4 * m, R = 6;
        std::vector<double> A( (size_t)(R * N), 0.0 );
        double rhs[6] = { 1, 1, 0, 0, 0, 0 };
        for( int i = 0; i < m; ++i )
        {
            for( int s = 0; s < 2; ++s )
            {
                const double t = s == 0 ? a : b;
                const int cl = 2 * i + s, cm = 2 * m + 2 * i + s;
                A[(size_t)(0 * N + cl)] = 1.0;
                A[(size_t)(1 * N + cm)] = 1.0;
                for( int k = 0; k < 3; ++k )
                {
                    A[(size_t)((2 + k) * N + cl)] =   p[3 * i + k] + t * u[3 * i + k];
                    A[(size_t)((2 + k) * N + cm)] = -(q[3 * i + k] + t * v[3 * i + k]);
                }
                A[(size_t)(5 * N + cl)] =   t * Tscale;
                A[(size_t)(5 * N + cm)] = -(t * Tscale);
            }
        }
        return LPFeasible( R, N, A.data(), rhs );
    }

    // CollisionFinder::FindMaximumSafeStepSize with the exact hull test.
    double SafeStepPair( int m, const double * p, const double * u, const double * q, const double * v, double tinit, double eps, double * tests )
    {
        double a = 0.0, b = tinit;
        // time scale of the space-time coordinates (it does not change whether the hulls meet)
        double cp[3] = {0,0,0}, cq[3] = {0,0,0}, r2 = 0;
        for( int i = 0; i < m; ++i ) { for( int k = 0; k < 3; ++k ) { cp[k] += p[3 * i + k] / m; cq[k] += q[3 * i + k] / m; } }
        for( int k = 0; k < 3; ++k ) { r2 += (cp[k] - cq[k]) * (cp[k] - cq[k]); }
        const double Tscale = std::sqrt( r2 ) / tinit + 1e-300;
        const int max_iter = 126;
        std::vector<double> stack( (size_t)max_iter + 2, 0.0 );
        int iter = 0, sp = 0;
        while( ( b - a > eps * b ) && ( iter < max_iter ) )
        {
            ++iter;
            if( tests ) { *tests += 1; }
            const bool meet = MovingHullsMeet( m, p, u, q, v, a, b, Tscale );
            if( !meet )
            {
                if( sp > 0 ) { a = b; b = stack[(size_t)sp--]; }
                else { return b; }
            }
            else { stack[(size_t)++sp] = b; b = 0.5 * ( a + b ); }
        }
        if( sp >= max_iter || iter >= max_iter ) { return a; }
        return ( a <= 0.0 ) ? b : a;
    }
}

extern "C"
{

// One pair (m = n+1 points each, rows p, u, q, v): the reference's bisection with the exact hull test.
double orc_safe_step_pair( int m, const double * p, const double * u, const double * q, const double * v, double tinit, double eps )
{
    return SafeStepPair( m, p, u, q, v, tinit, eps, nullptr );
}

// Do the space-time hulls over [a, b] meet?
int orc_moving_hulls_meet( int m, const double * p, const double * u, const double * q, const double * v, double a, double b )
{
    return MovingHullsMeet( m, p, u, q, v, a, b, 1.0 ) ? 1 : 0;
}

// Maximal safe step size over all pairs of non-adjacent simplices whose swept boxes over [0, t_max]
// meet (other pairs return t_max). X, U: V x 3 coordinates and update vectors; S: T x (n+1).
// stats: [0] candidate pairs, [1] hull tests.
double orc_max_safe_step( int n, LInt V, LInt T, const double * X, const double * U, const Int * S, double t_max, double eps, double * stats )
{
    (void)V;
    const int m = n + 1;
    std::vector<double> box( (size_t)(6 * T) );
    for( LInt i = 0; i < T; ++i )
    {
        double * bx = &box[(size_t)(6 * i)];
        for( int k = 0; k < 3; ++k ) { bx[k] = 1e300; bx[3 + k] = -1e300; }
        for( int l = 0; l < m; ++l )
        {
            const LInt w = S[m * i + l];
            for( int k = 0; k < 3; ++k )
            {
                const double x0 = X[3 * w + k], x1 = X[3 * w + k] + t_max * U[3 * w + k];
                bx[k] = std::min( bx[k], std::min( x0, x1 ) ); bx[3 + k] = std::max( bx[3 + k], std::max( x0, x1 ) );
            }
        }
    }
    double t = t_max;
    for( LInt i = 0; i < T; ++i )
    {
        for( LInt j = i + 1; j < T; ++j )
        {
            bool adjacent = false;
            for( int a = 0; a < m; ++a ) { for( int c = 0; c < m; ++c ) { adjacent = adjacent || ( S[m * i + a] == S[m * j + c] ); } }
            if( adjacent ) { continue; }
            const double * bi = &box[(size_t)(6 * i)], * bj = &box[(size_t)(6 * j)];
            bool overlap = true;
            for( int k = 0; k < 3; ++k ) { overlap = overlap && bi[k] <= bj[3 + k] && bj[k] <= bi[3 + k]; }
            if( !overlap ) { continue; }
            if( stats ) { stats[0] += 1; }
            double p[9], u[9], q[9], v[9];
            for( int l = 0; l < m; ++l )
            {
                for( int k = 0; k < 3; ++k )
                {
                    p[3 * l + k] = X[3 * (LInt)S[m * i + l] + k]; u[3 * l + k] = U[3 * (LInt)S[m * i + l] + k];
                    q[3 * l + k] = X[3 * (LInt)S[m * j + l] + k]; v[3 * l + k] = U[3 * (LInt)S[m * j + l] + k];
                }
            }
            t = std::min( t, SafeStepPair( m, p, u, q, v, t_max, eps, stats ? stats + 1 : nullptr ) );
        }
    }
    return t;
}

} // extern "C"
