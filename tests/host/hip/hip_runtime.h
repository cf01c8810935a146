// Host stand-in for <hip/hip_runtime.h>, used only to compile eik_stencil.h for the CPU in
// tests/host/replay_equiv.cpp: device qualifiers vanish and __ballot sees a single lane.
#pragma once
#define __device__
#define __forceinline__ inline
static inline unsigned long long __ballot(int x) { return x ? 1ull : 0ull; }
