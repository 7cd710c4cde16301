// shkz_stage.hpp -- dense staging and write-back rules of the Shiokaze modules, free of Shiokaze
// types so that the C++ host harness (shkz_harness.cpp) runs the same code as the modules.
//
// The modules densify the caller's sparse arrays with operator() (background / fill values off
// the active cells, include/shiokaze/array/array3.h:796-800), project through the C ABI, and
// write back what macpressuresolver3::project writes (src/projection/macpressuresolver3.cpp:
// 245-266):
//   * m_pressure: cleared, then set on every unknown of the system, zeros included (:245-248);
//   * velocity: only the faces the caller holds active are visited (velocity.parallel_actives,
//     :252); each keeps its projected value when the projection leaves it active (an open face,
//     or a closed boundary face with liquid inside, kept at 0), else it is switched off (:266).
#pragma once
#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace shkz_host {

// Runs fn(i) for i in [0, n) on every host thread, interleaved (the reference threads its loops
// over the CPU count, NUM_THREAD in wscript:75).
template <class Fn>
void parallel_for(int n, Fn fn) {
    unsigned nt = std::max(1u, std::thread::hardware_concurrency());
    nt = std::min<unsigned>(nt, (unsigned)std::max(n, 0));
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt);
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&fn, n, nt, t]() {
            for (int i = (int)t; i < n; i += (int)nt) fn(i);
        });
    for (std::thread &th : pool) th.join();
}

// Densifies a box (ni, nj, nk), z fastest, with value(i, j, k), over all host threads.
template <class T, class Get>
void densify(std::vector<T> &out, int ni, int nj, int nk, Get value) {
    out.resize((size_t)ni * nj * nk);
    T *o = out.data();
    parallel_for(ni, [&](int i) {
        for (int j = 0; j < nj; ++j) {
            T *row = o + ((size_t)i * nj + j) * nk;
            for (int k = 0; k < nk; ++k) row[k] = (T)value(i, j, k);
        }
    });
}

// The write-back of one face the caller holds active.
enum FaceAction { FACE_SKIP = 0, FACE_SET = 1, FACE_OFF = 2 };
inline FaceAction face_action(bool active_before, uint8_t active_after) {
    if (!active_before) return FACE_SKIP;  // parallel_actives does not visit it
    return active_after ? FACE_SET : FACE_OFF;
}

}  // namespace shkz_host
