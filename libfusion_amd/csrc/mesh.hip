// Marching-cubes mesh extraction over every allocated voxel block.
//
// Reference: src/voxel_hashing/scene_mesh.cu — BuildVertexArray (:14-248:
// read_sdf_list :70-106, interpolate_sdf :108-117, make_vertex :119-209,
// emitter :211-247), BuildVertexAndColourArray (:353-588), hosts
// create_mesh_vertex_only / _with_normal / _with_colour (:260-351,600-643);
// tables include/voxel_hashing/device_tables.h.
//
// MI355X design: one 256-thread workgroup per block.  The block's 9x9x9 corner
// lattice (its own 512 voxels plus the faces of the 7 +x/+y/+z neighbours) is
// staged in LDS after 8 hash lookups, instead of the reference's 8 hash walks
// per voxel.  Each lane then classifies two cubes; a workgroup prefix sum of
// the triangle counts reserves the output range with one atomic per block
// (the reference takes one global atomic per triangle).
#include "lf_common.h"
#include "mc_tables.h"

using namespace lf;

namespace {

__constant__ int c_edge[256] = LF_MC_EDGE_TABLE_DATA;
__constant__ signed char c_tri[256][16] = LF_MC_TRI_TABLE_DATA;

struct MeshDev {
    const unsigned long long *keys;
    const int *slot_block;
    const int *block_pos;
    const short *sdf;
    const float *weight;
    const unsigned char *rgb;
    unsigned int *counters;
    int num_buckets, num_entries, num_blocks;
    float voxel_size;
};

__global__ void k_mesh_begin(unsigned int *counters) { counters[LF_CNT_TRIANGLES] = 0u; }

LF_DEV float interp_edge(float v1, float v2) {  // interpolate_sdf (scene_mesh.cu:108-117)
    if ((double)fabsf(0.0f - v1) < 1e-6) return 0.0f;
    if ((double)fabsf(0.0f - v2) < 1e-6) return 1.0f;
    if ((double)fabsf(v1 - v2) < 1e-6) return 0.0f;
    return (0.0f - v1) / (v2 - v1);
}

LF_DEV int tri_count(int ci) {
    int n = 0;
    while (n < 15 && c_tri[ci][n] != -1) n += 3;
    return n / 3;
}

// mode 0: vertices, 1: + flat normals, 2: + corner colours.  kMixed: FP16
// sdf and 8-bit weight codes.
template <int kMode, bool kMixed>
__global__ __launch_bounds__(256) void k_mesh(MeshDev m, float *__restrict__ vtx, float *__restrict__ nrm,
                                              unsigned char *__restrict__ col, unsigned int max_tri) {
    __shared__ int s_nb[8];
    __shared__ float s_sdf[729];
    __shared__ unsigned char s_ok[729];
    __shared__ unsigned char s_col[kMode == 2 ? 729 * 3 : 1];
    __shared__ unsigned int s_wave_sum[4];
    __shared__ unsigned int s_base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned int nblk = min(m.counters[LF_CNT_ALLOCATED], (unsigned int)m.num_blocks);
    const float vs = m.voxel_size;
    for (unsigned int blk_i = blockIdx.x; blk_i < nblk; blk_i += gridDim.x) {
        const int4 P = reinterpret_cast<const int4 *>(m.block_pos)[blk_i];
        if (tid < 8)
            s_nb[tid] = tid == 0 ? (int)blk_i
                                 : find_block(m.keys, m.slot_block, P.x + (tid & 1), P.y + ((tid >> 1) & 1),
                                              P.z + ((tid >> 2) & 1), m.num_buckets, m.num_entries);
        __syncthreads();
        for (int i = tid; i < 729; i += 256) {
            const int x = i % 9, y = (i / 9) % 9, z = i / 81;
            const int nb = s_nb[(x >> 3) + 2 * (y >> 3) + 4 * (z >> 3)];
            unsigned char ok = 0;
            float s = 0.0f;
            if (nb >= 0) {
                const int l = lidx(x, y, z);
                const size_t vi = (size_t)nb * LF_BLOCK3 + l;
                // mixed: the weight code and the colour share one voxel word
                const unsigned int word = kMixed ? reinterpret_cast<const unsigned int *>(m.weight)[vi] : 0u;
                const float w = kMixed ? weight_value(word_code(word)) : m.weight[vi];
                ok = kMode == 2 ? ((double)w > 1e-3) : (w != 0.0f);
                s = kMixed ? half_bits_to_float(reinterpret_cast<const unsigned short *>(m.sdf)[vi])
                           : unpack_sdf(m.sdf[vi]);
                if (kMode == 2) {
                    if (kMixed) {
                        for (int k = 0; k < 3; ++k) s_col[3 * i + k] = (unsigned char)word_channel(word, k);
                    } else {
                        const unsigned char *c = m.rgb + (size_t)nb * (3 * LF_BLOCK3) + l;
                        s_col[3 * i] = c[0]; s_col[3 * i + 1] = c[LF_BLOCK3]; s_col[3 * i + 2] = c[2 * LF_BLOCK3];
                    }
                }
            }
            s_ok[i] = ok;
            s_sdf[i] = s;
        }
        __syncthreads();
        // classify this lane's two cubes
        int ci2[2];
        int cnt = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int v = tid + 256 * h;
            const int lx = v & 7, ly = (v >> 3) & 7, lz = v >> 6;
            int ci = 0;
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int cx = lx + ((k == 1 || k == 2 || k == 5 || k == 6) ? 1 : 0);
                const int cy = ly + ((k == 2 || k == 3 || k == 6 || k == 7) ? 1 : 0);
                const int cz = lz + (k >= 4 ? 1 : 0);
                const int li = cx + 9 * cy + 81 * cz;
                ok = ok && s_ok[li];
                if (s_sdf[li] < 0.0f) ci |= 1 << k;
            }
            if (!ok || c_edge[ci] == 0 || ci <= 0) ci = -1;
            ci2[h] = ci;
            if (ci > 0) cnt += tri_count(ci);
        }
        // workgroup exclusive scan of triangle counts
        unsigned int incl = (unsigned int)cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            unsigned int o = __shfl_up(incl, off);
            if (lane >= off) incl += o;
        }
        if (lane == 63) s_wave_sum[wave] = incl;
        __syncthreads();
        if (tid == 0) {
            unsigned int tot = s_wave_sum[0] + s_wave_sum[1] + s_wave_sum[2] + s_wave_sum[3];
            s_base = tot ? atomicAdd(&m.counters[LF_CNT_TRIANGLES], tot) : 0u;
        }
        __syncthreads();
        unsigned int id = s_base + incl - (unsigned int)cnt;
        for (int w = 0; w < wave; ++w) id += s_wave_sum[w];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ci = ci2[h];
            if (ci <= 0) continue;
            const int v = tid + 256 * h;
            const int lx = v & 7, ly = (v >> 3) & 7, lz = v >> 6;
            float s[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int cx = lx + ((k == 1 || k == 2 || k == 5 || k == 6) ? 1 : 0);
                const int cy = ly + ((k == 2 || k == 3 || k == 6 || k == 7) ? 1 : 0);
                const int cz = lz + (k >= 4 ? 1 : 0);
                s[k] = s_sdf[cx + 9 * cy + 81 * cz];
            }
            const float px = (float)(P.x * 8 + lx), py = (float)(P.y * 8 + ly), pz = (float)(P.z * 8 + lz);
            const int et = c_edge[ci];
            f3 va[12];
            if (et & 1) { float t = interp_edge(s[0], s[1]); va[0] = {px + t, py + 0.0f, pz + 0.0f}; }
            if (et & 2) { float t = interp_edge(s[1], s[2]); va[1] = {px + 1.0f, py + t, pz + 0.0f}; }
            if (et & 4) { float t = interp_edge(s[2], s[3]); va[2] = {px + (1.0f - t), py + 1.0f, pz + 0.0f}; }
            if (et & 8) { float t = interp_edge(s[3], s[0]); va[3] = {px + 0.0f, py + (1.0f - t), pz + 0.0f}; }
            if (et & 16) { float t = interp_edge(s[4], s[5]); va[4] = {px + t, py + 0.0f, pz + 1.0f}; }
            if (et & 32) { float t = interp_edge(s[5], s[6]); va[5] = {px + 1.0f, py + t, pz + 1.0f}; }
            if (et & 64) { float t = interp_edge(s[6], s[7]); va[6] = {px + (1.0f - t), py + 1.0f, pz + 1.0f}; }
            if (et & 128) { float t = interp_edge(s[7], s[4]); va[7] = {px + 0.0f, py + (1.0f - t), pz + 1.0f}; }
            if (et & 256) { float t = interp_edge(s[0], s[4]); va[8] = {px + 0.0f, py + 0.0f, pz + t}; }
            if (et & 512) { float t = interp_edge(s[1], s[5]); va[9] = {px + 1.0f, py + 0.0f, pz + t}; }
            if (et & 1024) { float t = interp_edge(s[2], s[6]); va[10] = {px + 1.0f, py + 1.0f, pz + t}; }
            if (et & 2048) { float t = interp_edge(s[3], s[7]); va[11] = {px + 0.0f, py + 1.0f, pz + t}; }
            for (int i = 0; i < 15 && c_tri[ci][i] != -1; i += 3, ++id) {
                if (id >= max_tri) continue;
                f3 t3[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int e = c_tri[ci][i + k];
                    t3[k] = mul(va[e], vs);
                    if (vtx) {
                        float *o = vtx + 9 * (size_t)id + 3 * k;
                        o[0] = t3[k].x; o[1] = t3[k].y; o[2] = t3[k].z;
                    }
                    if (kMode == 2 && col) {
                        // colour of the edge's first corner, corners 0-3 for edges 8-11
                        const int ce = e < 8 ? e : e - 8;
                        const int cx = lx + ((ce == 1 || ce == 2 || ce == 5 || ce == 6) ? 1 : 0);
                        const int cy = ly + ((ce == 2 || ce == 3 || ce == 6 || ce == 7) ? 1 : 0);
                        const int cz = lz + (ce >= 4 ? 1 : 0);
                        const int li = cx + 9 * cy + 81 * cz;
                        unsigned char *o = col + 9 * (size_t)id + 3 * k;
                        o[0] = s_col[3 * li]; o[1] = s_col[3 * li + 1]; o[2] = s_col[3 * li + 2];
                    }
                }
                if (kMode == 1 && nrm) {
                    const f3 n = normalised(cross(sub(t3[1], t3[0]), sub(t3[2], t3[0])));
                    float *o = nrm + 9 * (size_t)id;
#pragma unroll
                    for (int k = 0; k < 3; ++k) { o[3 * k] = n.x; o[3 * k + 1] = n.y; o[3 * k + 2] = n.z; }
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

extern "C" int lf_create_mesh(const lf_map_params *p, const lf_map_storage *s, float *vertex_out, float *normal_out,
                              unsigned char *colour_out, int max_triangles, unsigned int *triangles_out,
                              void *stream) {
    if (!p || !s || !triangles_out) return lf_set_error(LF_ERR_ARG, "null argument to create_mesh");
    if (max_triangles < 0) return lf_set_error(LF_ERR_ARG, "negative triangle capacity");
    if (normal_out && colour_out) return lf_set_error(LF_ERR_ARG, "ask for normals or colours, not both");
    if (int rc = lf_check_storage(s, p)) return rc;
    hipStream_t st = lf_stream(stream);
    MeshDev m;
    m.keys = s->keys; m.slot_block = s->slot_block; m.block_pos = s->block_pos; m.sdf = s->sdf;
    m.weight = s->weight; m.rgb = s->rgb; m.counters = s->counters;
    m.num_buckets = p->num_buckets; m.num_entries = p->num_entries; m.num_blocks = p->num_blocks;
    m.voxel_size = p->voxel_size;
    k_mesh_begin<<<1, 1, 0, st>>>(s->counters);
    LF_LAUNCH_CHECK();
    const int G = std::max(1, std::min(p->num_blocks, 8192));
    const unsigned int cap = (unsigned int)max_triangles;
    if (p->precision == LF_PRECISION_MIXED) {
        if (colour_out) k_mesh<2, true><<<G, 256, 0, st>>>(m, vertex_out, nullptr, colour_out, cap);
        else if (normal_out) k_mesh<1, true><<<G, 256, 0, st>>>(m, vertex_out, normal_out, nullptr, cap);
        else k_mesh<0, true><<<G, 256, 0, st>>>(m, vertex_out, nullptr, nullptr, cap);
    } else {
        if (colour_out) k_mesh<2, false><<<G, 256, 0, st>>>(m, vertex_out, nullptr, colour_out, cap);
        else if (normal_out) k_mesh<1, false><<<G, 256, 0, st>>>(m, vertex_out, normal_out, nullptr, cap);
        else k_mesh<0, false><<<G, 256, 0, st>>>(m, vertex_out, nullptr, nullptr, cap);
    }
    LF_LAUNCH_CHECK();
    unsigned int cnt = 0;
    LF_HIP_CHECK(hipMemcpyAsync(&cnt, s->counters + LF_CNT_TRIANGLES, sizeof(cnt), hipMemcpyDeviceToHost, st));
    LF_HIP_CHECK(hipStreamSynchronize(st));
    *triangles_out = std::min(cnt, cap);
    if (cnt > cap) return lf_set_error(LF_ERR_CAPACITY, "mesh has %u triangles, capacity %u", cnt, cap);
    return LF_OK;
}
