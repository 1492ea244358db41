// dbx_kernels.h -- host-callable launchers of the step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "dbx_device.h"

namespace dbx {

constexpr int kMaxBlock = 1024;

// Step kernels over the worlds order[worldOffset .. worldOffset + nworlds); lead
// selects the instantiation that profiles report for the lead (largest) tier.
struct WideBp;
// wide: the tier's worlds too big for one workgroup, whose first Collide pass
// runs over the whole device first
hipError_t launch_collide(const Dev &d, int cur, int worldOffset, int nworlds, int ldsBytes, int block, bool lead,
                          hipStream_t s, const WideBp *wide, int nwide);
size_t islands_scratch_bytes(int bodies, int contacts);
size_t solve_wave_lds_bytes(int bodies, int constraints, int levels, int islands);
hipError_t configure_kernels(int maxDynamicLds);
hipError_t launch_islands(const Dev &d, int cur, int ordered, int worldOffset, int nworlds, int ldsBytes, bool lead,
                          hipStream_t s);
// exact order: the island search of the tier's worlds too big for one workgroup
// (no joints), one island per wave
hipError_t launch_islands_wide(const Dev &d, int cur, const WideBp *wide, int nwide, hipStream_t s);
hipError_t launch_solve(const Dev &d, int cur, const StepParams &sp, int worldOffset, int nworlds, int block,
                        int ldsBytes, bool lead, hipStream_t s);
hipError_t launch_solve_wave(const Dev &d, int cur, const StepParams &sp, int worldOffset, int nworlds, int ldsBytes,
                             bool lead, hipStream_t s);
// exact order with kStepGrouped: the islands of the tier's big worlds (no joints)
// solved by groups of workgroups; the solve launches above skip those worlds
hipError_t launch_solve_groups(const Dev &d, int cur, const StepParams &sp, const WideBp *wide, int nwide,
                               int ldsBytes, bool lead, hipStream_t s);
size_t broadphase_scratch_bytes(int proxies);

// A world whose broadphase scratch does not fit in one workgroup's LDS: its pairs
// are found by a pass spread over the whole device (sort, sweep and pair sort in
// global memory) before k_broadphase adds them as contacts.
struct WideBp {
    int world;              // world index
    WorldDesc desc;         // its descriptor (fixture and pair ranges)
    int n;                  // proxies (fixtures)
    int P;                  // power of two >= n: sorted proxy keys
    int cap;                // the world's pair capacity
    int Pcap;               // power of two >= cap: sorted pair keys
    int nblk;               // 256-proxy blocks of the sweep count
    unsigned long long *keys; // P
    int *cur;               // n + 1 candidate offsets
    int *bsum;              // nblk block sums
    float4 *part;           // bounds partials, one per bounds workgroup
    int *ctr;               // [0] pairs found, [1] candidates, [2] sweep axis
    unsigned long long *pkeys; // Pcap
    // Collide's compaction and contact lists, and the continuous pass's lists of
    // the contacts made after Collide, over the whole device
    int nchunk;             // contact chunks of kWideCollideChunk (contact capacity)
    int2 *chunk;            // nchunk + 1: kept / touching contacts per chunk, then their offsets
    int *lcnt, *loff;       // bodyCount + 1: new contacts per body, then offsets
    int *lent;              // 2 * contact capacity: (2 * slot + side) per body
    // AddPair over the device: per pair its accept flag, per chunk of pairs the
    // accepted count, then its first contact slot
    int npchunk;            // pair chunks of kWideCollideChunk (pair capacity)
    int *pacc;              // pair capacity
    int2 *pchunk;           // npchunk + 1
    int tierFirst;          // 1 when the world is the first of its tier's launch order
};
constexpr int kWideCollideChunk = 1024;
// b.n, b.cap and b.desc set
size_t wide_scratch_bytes(const WideBp &b);
void wide_carve(WideBp &b, char *p);
hipError_t launch_broadphase(const Dev &d, int cur, int worldOffset, int nworlds, int ldsBytes, int block, bool lead,
                             hipStream_t s, const WideBp *wide = nullptr, int nwide = 0);
hipError_t launch_refit(const Dev &d, int world, hipStream_t s);
// b2World.SolveTOI (continuous physics) after the step's broadphase.
hipError_t launch_toi(const Dev &d, int cur, const StepParams &sp, int worldOffset, int nworlds, bool lead,
                      hipStream_t s, const WideBp *wide, int nwide);
hipError_t launch_bp_single(const float4 *fat, const int *moved, int n, char *scratch, int2 *out, int cap, int *total,
                            hipStream_t s);
hipError_t launch_collide_pairs(const Shape *sa, const float *xa, const Shape *sb, const float *xb, int n, float *out,
                                hipStream_t s);
hipError_t launch_max_penetration(const Dev &d, int cur, float *out, hipStream_t s);

} // namespace dbx
