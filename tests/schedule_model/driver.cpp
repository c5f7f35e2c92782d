/*
 * driver.cpp — runs the engine's host code (under fake_hip.cpp) through a
 * multi-stream schedule and leaves the trace for the checker.
 *
 *   model_driver STREAMS GROUPS STEPS PATTERN [pipeline] [queues] [level_min] [fused_step]
 *
 * Streams of four frame sizes (preset 3, one seed per stream) are opened;
 * each step submits one frame of every stream as GROUPS batches, stream s
 * in group s % GROUPS (PATTERN "alt"), or regrouped every step (PATTERN
 * "rotate": stream s in group (s + step) % GROUPS), or GROUPS batches on
 * even steps and one batch on odd ones ("switch"). Frame 0 is a KEY frame,
 * later steps cycle through the inter frames. From step fused_step on (if
 * given) the engine runs fused. The engine is synchronised and every stream
 * read back at the end.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "cvp9_hip/engine.h"
#include "cvp9_hip/synth.h"

void cvp9_model_note(const char *text);

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s STREAMS GROUPS STEPS alt|rotate|switch [pipeline] [queues] [level_min] [fused_step]\n",
                argv[0]);
        return 2;
    }
    const int ns = atoi(argv[1]), groups = atoi(argv[2]), steps = atoi(argv[3]);
    const std::string pattern = argv[4];
    const int pipeline = argc > 5 ? atoi(argv[5]) : 1;
    const int queues = argc > 6 ? atoi(argv[6]) : 2;
    const int level_min = argc > 7 ? atoi(argv[7]) : 0;
    const int fused_step = argc > 8 ? atoi(argv[8]) : -1;
    const int F = 4; /* frames per stream: a KEY frame and three inter frames */
    const unsigned sizes[4][2] = {{64, 48}, {96, 64}, {136, 72}, {160, 96}};

    cvp9_engine_t *e = nullptr;
    if (cvp9_engine_create(0, &e) != CVP9_OK) return 1;
    cvp9_engine_set_pipeline(e, pipeline);
    cvp9_engine_set_queues(e, queues);
    cvp9_tune_t t;
    cvp9_engine_get_tune(e, &t);
    t.intra_level_min = (uint32_t)level_min;
    cvp9_engine_set_tune(e, &t);

    std::vector<int> sid(ns);
    std::vector<std::vector<cvp9_prep_t *>> prep(ns);
    for (int s = 0; s < ns; s++) {
        if (cvp9_engine_stream_open(e, &sid[s]) != CVP9_OK) return 1;
        cvp9_synth_params_t p;
        cvp9_synth_preset(&p, 3, sizes[s % 4][0], sizes[s % 4][1], 7000 + s);
        for (int i = 0; i < F; i++) {
            vp9_parsed_frame_t *pf = cvp9_synth_frame(&p, (uint32_t)i);
            cvp9_prep_t *pp = nullptr;
            if (!pf || cvp9_engine_prepare(e, pf, &pp) != CVP9_OK) return 1;
            cvp9_synth_free(pf);
            prep[s].push_back(pp);
        }
    }
    char note[128];
    for (int k = 0; k < steps; k++) {
        if (fused_step >= 0 && k == fused_step && cvp9_engine_set_fused(e, 1) != CVP9_OK) return 1;
        const int idx = k < F ? k : 1 + (k - 1) % (F - 1);
        const int G = pattern == "switch" && (k & 1) ? 1 : groups;
        for (int g = 0; g < G; g++) {
            std::vector<int> ids;
            std::vector<cvp9_prep_t *> pp;
            for (int s = 0; s < ns; s++) {
                const int grp = pattern == "rotate" ? (s + k) % G : s % G;
                if (grp != g) continue;
                ids.push_back(sid[s]);
                pp.push_back(prep[s][idx]);
            }
            if (ids.empty()) continue;
            snprintf(note, sizeof note, "batch step %d group %d streams %zu", k, g, ids.size());
            cvp9_model_note(note);
            if (cvp9_engine_run(e, (int)ids.size(), ids.data(), pp.data(), nullptr) != CVP9_OK) {
                fprintf(stderr, "run failed at step %d group %d\n", k, g);
                return 1;
            }
        }
    }
    cvp9_model_note("sync");
    if (cvp9_engine_sync(e) != CVP9_OK) return 1;
    std::vector<uint8_t> buf(160 * 96 * 2);
    for (int s = 0; s < ns; s++) {
        uint32_t w = 0, h = 0;
        snprintf(note, sizeof note, "read stream %d", s);
        cvp9_model_note(note);
        if (cvp9_engine_read(e, sid[s], buf.data(), buf.size(), &w, &h) != CVP9_OK) return 1;
    }
    for (int s = 0; s < ns; s++)
        for (cvp9_prep_t *pp : prep[s]) cvp9_engine_prep_free(e, pp);
    cvp9_engine_destroy(e);
    return 0;
}
