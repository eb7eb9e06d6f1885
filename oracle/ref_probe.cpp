/*
 * ref_probe.cpp - TEST INFRASTRUCTURE ONLY.  A driver around the unmodified
 * reference source src/CPU/ndt_mapping/kernel.cpp, built by oracle/Makefile into
 * oracle/_ref/ndt_ref_probe when the reference tree is present.  It calls the
 * reference's own member functions on inputs written by tests/golden/make_golden.py
 * and writes their raw float64/float32 outputs, which become the golden fixtures
 * that pin the oracle function by function.
 *
 * The reference file is compiled in this translation unit with its private and
 * protected members opened up by the preprocessor; its text is not modified.
 *
 * Input/output files: little-endian raw arrays, layouts documented per command.
 */
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#define private public
#define protected public
#include "kernel.cpp"
#undef private
#undef protected

static std::vector<char> slurp(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path);
        std::exit(2);
    }
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

struct reader {
    std::vector<char> buf;
    size_t pos = 0;
    template <class T> T get() {
        T v;
        std::memcpy(&v, buf.data() + pos, sizeof(T));
        pos += sizeof(T);
        return v;
    }
};

struct writer {
    FILE* f;
    template <class T> void put(T v) { std::fwrite(&v, sizeof(T), 1, f); }
};

/* one DAPHNE testcase: guess float[16], scan (int32 n, n*float[4]), map (int32 m, m*float[4]) */
static void read_case(reader& r, Matrix4f& guess, PointCloud& scan, PointCloud& map) {
    for (int h = 0; h < 4; h++)
        for (int w = 0; w < 4; w++) guess.data[h][w] = r.get<float>();
    int32_t n = r.get<int32_t>();
    scan.resize(n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) scan[i].data[k] = r.get<float>();
    int32_t m = r.get<int32_t>();
    map.resize(m);
    for (int i = 0; i < m; i++)
        for (int k = 0; k < 4; k++) map[i].data[k] = r.get<float>();
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <cmd> <in> <out>\n", argv[0]);
        return 2;
    }
    std::string cmd = argv[1];
    reader r;
    r.buf = slurp(argv[2]);
    writer w;
    w.f = std::fopen(argv[3], "wb");
    if (!w.f) return 2;
    ndt_mapping& k = a; /* the reference's global instance */

    if (cmd == "case") {
        /* in: one DAPHNE testcase (no count header).
         * out: grid {min float[4], max float[4], dims int32[3], ncells int64,
         *            nocc int64, nocc x {idx int64, n int32, mean f64[3], icov f64[9]}},
         *      pass0 {p f64[6], score f64, grad f64[6], hess f64[36]},
         *      radius {K int32, K x {count int32, up to 27 x {mean f64[3]}}} for the
         *              first min(n, 64) guess-transformed points,
         *      result {final float[16], converged uint8, iterations int32} */
        Matrix4f guess;
        PointCloud scan, map;
        read_case(r, guess, scan, map);
        CallbackResult res = k.partial_points_callback(scan, guess, map);
        int iterations = k.nr_iterations_;
        /* the callback left the grid of this map and the Gauss constants in place */
        for (int i = 0; i < 4; i++) w.put<float>(k.minVoxel.data[i]);
        for (int i = 0; i < 4; i++) w.put<float>(k.maxVoxel.data[i]);
        for (int i = 0; i < 3; i++) w.put<int32_t>(k.voxelDimension[i]);
        w.put<int64_t>((int64_t)k.target_cells_.size());
        int64_t nocc = 0;
        for (size_t i = 0; i < k.target_cells_.size(); i++)
            if (k.target_cells_[i].numberPoints > 0) nocc++;
        w.put<int64_t>(nocc);
        for (size_t i = 0; i < k.target_cells_.size(); i++) {
            const Voxel& v = k.target_cells_[i];
            if (v.numberPoints <= 0) continue;
            w.put<int64_t>((int64_t)i);
            w.put<int32_t>(v.numberPoints);
            for (int c = 0; c < 3; c++) w.put<double>(v.mean[c]);
            for (int rr = 0; rr < 3; rr++)
                for (int c = 0; c < 3; c++) w.put<double>(v.invCovariance.data[rr][c]);
        }
        /* first derivative pass, as computeTransformation sets it up (:1187-1213) */
        PointCloud output(scan);
        for (size_t i = 0; i < output.size(); i++) output[i].data[3] = 1.0;
        k.input_ = &scan;
        transformPointCloud(output, output, guess);
        std::memset(k.point_gradient_.data, 0, sizeof(double) * 3 * 6);
        k.point_gradient_.data[0][0] = 1.0;
        k.point_gradient_.data[1][1] = 1.0;
        k.point_gradient_.data[2][2] = 1.0;
        std::memset(k.point_hessian_.data, 0, sizeof(double) * 18 * 6);
        Vec6 p, grad;
        Mat66 hess;
        p[0] = guess.data[0][3];
        p[1] = guess.data[1][3];
        p[2] = guess.data[2][3];
        Vec3 ea;
        k.eulerAngles(guess, ea);
        p[3] = ea[0];
        p[4] = ea[1];
        p[5] = ea[2];
        double score = k.computeDerivatives(grad, hess, output, p);
        for (int i = 0; i < 6; i++) w.put<double>(p[i]);
        w.put<double>(score);
        for (int i = 0; i < 6; i++) w.put<double>(grad[i]);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) w.put<double>(hess.data[i][j]);
        int K = (int)std::min<size_t>(output.size(), 64);
        w.put<int32_t>(K);
        for (int i = 0; i < K; i++) {
            std::vector<Voxel> nb;
            std::vector<float> dist;
            int cnt = k.voxelRadiusSearch(k.target_cells_, output[i], k.resolution_, nb, dist);
            w.put<int32_t>(cnt);
            for (int q = 0; q < cnt && q < 27; q++)
                for (int c = 0; c < 3; c++) w.put<double>(nb[q].mean[c]);
        }
        for (int h = 0; h < 4; h++)
            for (int ww = 0; ww < 4; ww++) w.put<float>(res.final_transformation.data[h][ww]);
        w.put<uint8_t>(res.converged ? 1 : 0);
        w.put<int32_t>(iterations);
        w.put<double>(k.gauss_d1_);
        w.put<double>(k.gauss_d2_);
    } else if (cmd == "math") {
        /* in: int32 K, then K records {p f64[6], x f64[3], A f64[36], b f64[6],
         *     mt f64[9], xt f64[3], cinv f64[9], M float[16]}
         * out per record: jang f64[8*3], hang f64[15*3], pg f64[18], ph f64[108],
         *     solve f64[6], trial f64, upd {ret int32, a_l f_l g_l a_u f_u g_u f64},
         *     update {score_inc f64, grad f64[6], hess f64[36]},
         *     build float[16], euler f64[3] */
        int32_t K = r.get<int32_t>();
        k.gauss_d1_ = -2.2172252440428890;
        k.gauss_d2_ = 0.43312300470355464;
        for (int rec = 0; rec < K; rec++) {
            Vec6 p;
            Vec3 x;
            Mat66 A;
            Vec6 b;
            double mt[9];
            Vec3 xt;
            Mat33 cinv;
            Matrix4f M;
            for (int i = 0; i < 6; i++) p[i] = r.get<double>();
            for (int i = 0; i < 3; i++) x[i] = r.get<double>();
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) A.data[i][j] = r.get<double>();
            for (int i = 0; i < 6; i++) b[i] = r.get<double>();
            for (int i = 0; i < 9; i++) mt[i] = r.get<double>();
            for (int i = 0; i < 3; i++) xt[i] = r.get<double>();
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) cinv.data[i][j] = r.get<double>();
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) M.data[i][j] = r.get<float>();
            k.computeAngleDerivatives(p);
            Vec3* jt[8] = {&k.j_ang_a_, &k.j_ang_b_, &k.j_ang_c_, &k.j_ang_d_,
                           &k.j_ang_e_, &k.j_ang_f_, &k.j_ang_g_, &k.j_ang_h_};
            Vec3* ht[15] = {&k.h_ang_a2_, &k.h_ang_a3_, &k.h_ang_b2_, &k.h_ang_b3_, &k.h_ang_c2_,
                            &k.h_ang_c3_, &k.h_ang_d1_, &k.h_ang_d2_, &k.h_ang_d3_, &k.h_ang_e1_,
                            &k.h_ang_e2_, &k.h_ang_e3_, &k.h_ang_f1_, &k.h_ang_f2_, &k.h_ang_f3_};
            for (int i = 0; i < 8; i++)
                for (int c = 0; c < 3; c++) w.put<double>((*jt[i])[c]);
            for (int i = 0; i < 15; i++)
                for (int c = 0; c < 3; c++) w.put<double>((*ht[i])[c]);
            std::memset(k.point_gradient_.data, 0, sizeof(double) * 3 * 6);
            k.point_gradient_.data[0][0] = 1.0;
            k.point_gradient_.data[1][1] = 1.0;
            k.point_gradient_.data[2][2] = 1.0;
            std::memset(k.point_hessian_.data, 0, sizeof(double) * 18 * 6);
            k.computePointDerivatives(x);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 6; j++) w.put<double>(k.point_gradient_.data[i][j]);
            for (int i = 0; i < 18; i++)
                for (int j = 0; j < 6; j++) w.put<double>(k.point_hessian_.data[i][j]);
            Vec6 sol;
            solve(sol, A, b);
            for (int i = 0; i < 6; i++) w.put<double>(sol[i]);
            double trial = k.trialValueSelectionMT(mt[0], mt[1], mt[2], mt[3], mt[4], mt[5], mt[6], mt[7], mt[8]);
            w.put<double>(trial);
            double a_l = mt[0], f_l = mt[1], g_l = mt[2], a_u = mt[3], f_u = mt[4], g_u = mt[5];
            bool ret = k.updateIntervalMT(a_l, f_l, g_l, a_u, f_u, g_u, mt[6], mt[7], mt[8]);
            w.put<int32_t>(ret ? 1 : 0);
            w.put<double>(a_l); w.put<double>(f_l); w.put<double>(g_l);
            w.put<double>(a_u); w.put<double>(f_u); w.put<double>(g_u);
            Vec6 g6 = {0, 0, 0, 0, 0, 0};
            Mat66 h6;
            std::memset(&h6, 0, sizeof(h6));
            double inc = k.updateDerivatives(g6, h6, xt, cinv, true);
            w.put<double>(inc);
            for (int i = 0; i < 6; i++) w.put<double>(g6[i]);
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) w.put<double>(h6.data[i][j]);
            Matrix4f B = M;
            k.buildTransformationMatrix(B, p);
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) w.put<float>(B.data[i][j]);
            Vec3 ea;
            k.eulerAngles(M, ea);
            for (int i = 0; i < 3; i++) w.put<double>(ea[i]);
        }
    } else {
        std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
        return 2;
    }
    std::fclose(w.f);
    return 0;
}
