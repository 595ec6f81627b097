/*
 * atrc_host.h — C surface of libatrc_host.so, the C++ host that reads Atrc
 * JSON (scene + rendering sessions, src/cli/src/main.cpp:14-72) and renders
 * through the "gpu_pt" renderer.  Used by language bindings and tests; the
 * CLI (atrc_cli -d scene.json | -s "<json>") is the same code as an executable.
 */
#ifndef ATRC_HOST_H
#define ATRC_HOST_H

#include "atrc_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct atrc_host_session {
    int32_t width, height;
    float eps;
    float filter_radius;
    agx_camera_desc camera;
    int32_t session_count;
    int32_t filter;            /* AGX_FILTER_* */
    float filter_alpha;        /* Gaussian alpha */
} atrc_host_session;

/* Parse JSON text ({"scene": ..., "rendering": ...}) and create the scene and
 * every render session; scene_dir replaces ${scene-directory}.  NULL on error. */
void *atrc_host_load(const char *json_text, const char *scene_dir);
/* Descriptor of the loaded scene (valid until atrc_host_free). */
const agx_scene_desc *atrc_host_scene_desc(void *handle);
int  atrc_host_session_info(void *handle, int index, atrc_host_session *out);
/* RenderSession::execute (render + post-processors); outputs as agx_render's, may be NULL. */
int  atrc_host_execute(void *handle, int index, float *rgb, float *albedo, float *normal, float *denoise);
void atrc_host_free(void *handle);
/* Write a ".bm" binary mesh (uint64 count, then 3 agx_vertex per triangle). */
int  atrc_host_write_bin_mesh(const char *filename, const agx_vertex *tris, uint32_t tri_count);
const char *atrc_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ATRC_HOST_H */
