/*
 * rtc8_oracle_nrc.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) of the NRC data path
 * (renderer/gpu_kernels/nrc_set_up_kernels.cu): prepareNRC, propagateRadianceValues,
 * shuffleTrainingData, accumulateInferredRadianceValues.  Operates on host copies of the buffers in
 * rtc8x_frame_params.nrc; FP32 arithmetic in the reference's expression order.
 */
#include "rtc8_oracle.h"

#include <math.h>
#include <string.h>

static void rgb_load(const float* p, float c[3]) { c[0] = p[0]; c[1] = p[1]; c[2] = p[2]; }

/* prepareNRC (nrc_set_up_kernels.cu:4-44); the new size is clamped to >= 1 pixel. */
void oracle_nrc_prepare(const rtc8x_frame_params* p, int32_t is_new_sequence)
{
    const rtc8x_nrc_buffers* nb = &p->nrc;
    uint32_t buf = nb->nrc_buffer_index, prev = (buf + 1) % 2;
    int32_t nw, nh;
    if (is_new_sequence) {
        nw = (int32_t)p->image_width / 4;
        nh = (int32_t)p->image_height / 4;
    } else {
        float r = sqrtf((float)nb->num_training_data[prev] / (float)RTC8X_NUM_TRAINING_DATA_PER_FRAME);
        int32_t cw = nb->train_image_size[2 * prev], ch = nb->train_image_size[2 * prev + 1];
        nw = (int32_t)fminf((float)cw / r, (float)p->image_width);
        nh = (int32_t)fminf((float)ch / r, (float)p->image_height);
    }
    if (nw < 1) nw = 1;
    if (nh < 1) nh = 1;
    nb->train_image_size[2 * buf] = nw;
    nb->train_image_size[2 * buf + 1] = nh;
    nb->num_training_data[buf] = 0;
    for (int c = 0; c < 3; ++c) {
        nb->target_min_max[6 * buf + c] = INFINITY;
        nb->target_min_max[6 * buf + 3 + c] = -INFINITY;
        nb->target_avg[3 * buf + c] = 0.0f;
    }
    for (uint32_t i = 0; i < nb->max_num_training_suffixes; ++i) {
        nb->train_suffix_terminals[i].prev_vertex_index = RTC8X_INVALID_VERTEX_INDEX;
        nb->train_suffix_terminals[i].has_query = 0;
        nb->train_suffix_terminals[i].path_length = 0;
        nb->train_suffix_terminals[i].is_unbiased = 0;
    }
}

/* propagateRadianceValues (nrc_set_up_kernels.cu:46-86) */
void oracle_nrc_propagate(const rtc8x_frame_params* p)
{
    const rtc8x_nrc_buffers* nb = &p->nrc;
    uint32_t buf = nb->nrc_buffer_index;
    uint32_t n = (uint32_t)nb->train_image_size[2 * buf] * (uint32_t)nb->train_image_size[2 * buf + 1];
    for (uint32_t i = 0; i < n; ++i) {
        const rtc8x_training_suffix_terminal* ti = &nb->train_suffix_terminals[i];
        if (ti->prev_vertex_index == RTC8X_INVALID_VERTEX_INDEX) continue;
        float c[3] = {0.0f, 0.0f, 0.0f};
        if (ti->has_query) {
            rgb_load(nb->inferred_radiance + 3 * (size_t)i, c);
            for (int k = 0; k < 3; ++k) c[k] = fmaxf(c[k], 0.0f);
            if (nb->radiance_scale > 0.0f) {
                float r = 1.0f / nb->radiance_scale;
                for (int k = 0; k < 3; ++k) c[k] = c[k] * r;
            }
            const rtc8x_radiance_query* q = &nb->inference_queries[i];
            for (int k = 0; k < 3; ++k) c[k] = c[k] * (q->diffuse_reflectance[k] + q->specular_reflectance[k]);
        }
        uint32_t idx = ti->prev_vertex_index;
        while (idx != RTC8X_INVALID_VERTEX_INDEX) {
            const rtc8x_training_vertex_info* v = &nb->train_vertex_info[idx];
            float* tgt = nb->train_targets[0] + 3 * (size_t)idx;
            const rtc8x_radiance_query* q = &nb->train_queries[0][idx];
            for (int k = 0; k < 3; ++k) {
                float indirect = v->local_throughput[k] * c[k];
                c[k] = tgt[k] + indirect;
                float ref = q->diffuse_reflectance[k] + q->specular_reflectance[k];
                tgt[k] = ref != 0.0f ? c[k] / ref : 0.0f;
            }
            idx = v->prev_vertex_index;
        }
    }
}

static int query_valid(const rtc8x_radiance_query* q)
{
    const float* f = (const float*)q;
    for (int k = 0; k < 14; ++k)
        if (!isfinite(f[k])) return 0;
    return 1;
}

/* shuffleTrainingData (nrc_set_up_kernels.cu:88-174).  Statistics are reduced in the fixed order the
 * library uses: 256-record blocks by a halving tree, then the 256 block results by the same tree. */
void oracle_nrc_shuffle(const rtc8x_frame_params* p)
{
    const rtc8x_nrc_buffers* nb = &p->nrc;
    uint32_t buf = nb->nrc_buffer_index;
    uint32_t num = nb->num_training_data[buf];
    enum { B = 256, NB = RTC8X_NUM_TRAINING_DATA_PER_FRAME / 256 };
    static float part[NB][9];
    for (uint32_t blk = 0; blk < NB; ++blk) {
        float smin[3][B], smax[3][B], ssum[3][B];
        for (uint32_t t = 0; t < B; ++t) {
            uint32_t i = blk * B + t;
            if (num > 0) {
                uint32_t state = nb->data_shufflers[i];
                state = ((state * 1103515245u) + 12345u) % (1u << 31);
                nb->data_shufflers[i] = state;
                uint32_t dst = state % RTC8X_NUM_TRAINING_DATA_PER_FRAME;
                uint32_t src = i % num;
                rtc8x_radiance_query q = nb->train_queries[0][src];
                float tg[3];
                rgb_load(nb->train_targets[0] + 3 * (size_t)src, tg);
                if (!query_valid(&q)) memset(&q, 0, sizeof(q));
                int ok = 1;
                for (int k = 0; k < 3; ++k) ok = ok && tg[k] >= 0.0f && isfinite(tg[k]);
                if (!ok) tg[0] = tg[1] = tg[2] = 0.0f;
                for (int k = 0; k < 3; ++k) {
                    smin[k][t] = tg[k];
                    smax[k][t] = tg[k];
                    ssum[k][t] = tg[k] * (1.0f / (float)RTC8X_NUM_TRAINING_DATA_PER_FRAME);
                }
                float sc[3] = {tg[0], tg[1], tg[2]};
                if (nb->radiance_scale > 0.0f)
                    for (int k = 0; k < 3; ++k) sc[k] = sc[k] * nb->radiance_scale;
                for (int k = 0; k < 3; ++k) sc[k] = fminf(sc[k], 1e6f);
                nb->train_queries[1][dst] = q;
                memcpy(nb->train_targets[1] + 3 * (size_t)dst, sc, sizeof(sc));
            } else {
                memset(&nb->train_queries[1][i], 0, sizeof(rtc8x_radiance_query));
                memset(nb->train_targets[1] + 3 * (size_t)i, 0, 3 * sizeof(float));
                for (int k = 0; k < 3; ++k) {
                    smin[k][t] = INFINITY;
                    smax[k][t] = -INFINITY;
                    ssum[k][t] = 0.0f;
                }
            }
        }
        for (uint32_t s = B / 2; s > 0; s >>= 1)
            for (uint32_t t = 0; t < s; ++t)
                for (int k = 0; k < 3; ++k) {
                    smin[k][t] = fminf(smin[k][t], smin[k][t + s]);
                    smax[k][t] = fmaxf(smax[k][t], smax[k][t + s]);
                    ssum[k][t] = ssum[k][t] + ssum[k][t + s];
                }
        for (int k = 0; k < 3; ++k) {
            part[blk][k] = smin[k][0];
            part[blk][3 + k] = smax[k][0];
            part[blk][6 + k] = ssum[k][0];
        }
    }
    if (num == 0) return;
    float smin[3][NB], smax[3][NB], ssum[3][NB];
    for (uint32_t t = 0; t < NB; ++t)
        for (int k = 0; k < 3; ++k) {
            smin[k][t] = part[t][k];
            smax[k][t] = part[t][3 + k];
            ssum[k][t] = part[t][6 + k];
        }
    for (uint32_t s = NB / 2; s > 0; s >>= 1)
        for (uint32_t t = 0; t < s; ++t)
            for (int k = 0; k < 3; ++k) {
                smin[k][t] = fminf(smin[k][t], smin[k][t + s]);
                smax[k][t] = fmaxf(smax[k][t], smax[k][t + s]);
                ssum[k][t] = ssum[k][t] + ssum[k][t + s];
            }
    for (int k = 0; k < 3; ++k) {
        nb->target_min_max[6 * buf + k] = smin[k][0];
        nb->target_min_max[6 * buf + 3 + k] = smax[k][0];
        nb->target_avg[3 * buf + k] = ssum[k][0];
    }
}

/* accumulateInferredRadianceValues (nrc_set_up_kernels.cu:176-219), FP64 accumulation, once per sample
 * slice of the launch (nrc.num_samples), in sample order. */
void oracle_nrc_accumulate(const rtc8x_frame_params* p)
{
    const rtc8x_nrc_buffers* nb = &p->nrc;
    uint32_t n = p->image_width * p->image_height;
    uint32_t samples = nb->num_samples ? nb->num_samples : 1u;
    size_t q_stride = ((size_t)n + 127u) & ~(size_t)127u;
    for (uint32_t s = 0; s < samples; ++s)
    for (uint32_t px = 0; px < n; ++px) {
        size_t i = (size_t)s * n + px, qi = (size_t)s * q_stride + px;
        const rtc8x_terminal_info* ti = &nb->inference_terminals[i];
        float direct[3], rad[3] = {0.0f, 0.0f, 0.0f};
        rgb_load(nb->per_frame_contribution + 3 * i, direct);
        if (ti->has_query & 1u) {
            rgb_load(nb->inferred_radiance + 3 * qi, rad);
            for (int k = 0; k < 3; ++k) rad[k] = fmaxf(rad[k], 0.0f);
            if (nb->radiance_scale > 0.0f) {
                float r = 1.0f / nb->radiance_scale;
                for (int k = 0; k < 3; ++k) rad[k] = rad[k] * r;
            }
            const rtc8x_radiance_query* q = &nb->inference_queries[qi];
            for (int k = 0; k < 3; ++k) rad[k] = rad[k] * (q->diffuse_reflectance[k] + q->specular_reflectance[k]);
        }
        for (int k = 0; k < 3; ++k) {
            float indirect = ti->throughput[k] * rad[k];
            float c = direct[k] + indirect;
            p->accum[3 * (size_t)px + k] += (double)c;
        }
    }
}
