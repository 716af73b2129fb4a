// ORACLE -- test infrastructure only.
// A C-callable driver around the reference's instance-file layer, gpu/multibit_bias_20261008/src/jmat.hpp (loaders
// from_jint8 :86-103, from_gset :65-84, from_k2000 :57-63 over sca::load_graph fpga/v80_sca/src/sca_ref.hpp:31-49,
// write_jint8 :105-113, load_bias/write_bias :164-186; finish :35-55; scorers scalar_cut/energy/field_mismatches
// :130-162), compiled from the reference sources where they lie (oracle/Makefile, development container only). It
// checks the library's own readers, writers and scorers in tests/test_instance_io.py; it is never part of the
// product library.
// Matrices cross this boundary in the reference's layout: NP x NP int8 (NP = 2048), zero padding.
#include "jmat.hpp"

#include <cstring>
#include <exception>
#include <string>

namespace {
thread_local std::string g_msg;

int caught(const std::exception& e) {
    g_msg = e.what();
    return -1;
}

void put(const jmat::IntJ& m, int8_t* dense, int* n, int64_t* st) {
    std::memcpy(dense, m.dense.data(), m.dense.size());
    *n = m.N;
    st[0] = m.sumw; st[1] = m.maxabs; st[2] = m.hmax; st[3] = int64_t(m.row_ptr.back());
}

jmat::IntJ make(int n, const int8_t* dense) {
    jmat::IntJ m;
    m.N = n;
    std::memcpy(m.dense.data(), dense, m.dense.size());
    jmat::finish(m);
    return m;
}
}  // namespace

extern "C" {

const char* ref_io_error() { return g_msg.c_str(); }

// kind: 0 jint8, 1 graph (K2000 SNOWGPU1), 2 gset. st: sumw, maxabs, hmax, nnz. mean_row_sq separately.
int ref_io_load(int kind, const char* path, int8_t* dense, int* n, int64_t* st, double* mean_row_sq) {
    try {
        const jmat::IntJ m = jmat::load(kind == 0 ? "jint8" : kind == 1 ? "graph" : "gset", path);
        put(m, dense, n, st);
        *mean_row_sq = m.mean_row_sq;
        return 0;
    } catch (const std::exception& e) { return caught(e); }
}

int ref_io_write_jint8(int n, const int8_t* dense, const char* path) {
    try {
        jmat::write_jint8(make(n, dense), path);
        return 0;
    } catch (const std::exception& e) { return caught(e); }
}

int ref_io_load_bias(const char* path, int n, int32_t* b) {
    try {
        const std::vector<int32_t> v = jmat::load_bias(path, n);
        std::memcpy(b, v.data(), sizeof(int32_t) * v.size());
        return 0;
    } catch (const std::exception& e) { return caught(e); }
}

int ref_io_write_bias(int n, const int32_t* b, const char* path) {
    try {
        jmat::write_bias(std::vector<int32_t>(b, b + n), n, path);
        return 0;
    } catch (const std::exception& e) { return caught(e); }
}

// bias: NP entries; bits: NP/64 words; hdev: NP entries (only the first n are compared). out: cut, energy, mismatches.
int ref_io_score(int n, const int8_t* dense, const int32_t* bias, const uint64_t* bits, const int32_t* hdev, int64_t* out) {
    try {
        const jmat::IntJ m = make(n, dense);
        out[0] = jmat::scalar_cut(m, bits);
        out[1] = jmat::energy(m, bias, bits);
        out[2] = hdev ? jmat::field_mismatches(m, bias, bits, hdev) : 0;
        return 0;
    } catch (const std::exception& e) { return caught(e); }
}

}  // extern "C"
