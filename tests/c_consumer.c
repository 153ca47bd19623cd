/* A C99 caller of the C ABI, built and run by tests/test_c_consumer.py with
 * the system C compiler.  It uses only the entry points that need no GPU:
 * descriptor defaults, validation, device-byte estimates, MRC round trip,
 * the thickness edge search, and the error string on failure.  Prints "ok"
 * and exits 0 when every check holds; otherwise names the failed check. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "aretomo_files.h"
#include "aretomo_recon.h"

#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            printf("FAILED %s:%d %s\n", __FILE__, __LINE__, #cond);   \
            return 1;                                                 \
        }                                                             \
    } while (0)

int main(int argc, char** argv)
{
    enum { N = 11, NX = 24, NY = 6, NZ = 5 };
    float tilts[N];
    atr_series_desc d;
    size_t small = 0, large = 0;
    char path[4096];
    float vol[NX * NY * NZ], back[NX * NY * NZ];
    atr_mrc_info info;
    float profile[40];
    atr_thickness th;
    int i;

    if (argc < 2) return 2;
    CHECK(atr_version() != NULL && strlen(atr_version()) > 0);

    /* reference defaults (CTomoSart / CDoSartRecon) */
    atr_series_desc_init(&d);
    CHECK(d.method == ATR_METHOD_SART);
    CHECK(d.sart_iters == 20 && d.positivity == 1);
    CHECK(d.subset_order == ATR_SUBSETS_CONTIGUOUS && d.storage == ATR_STORAGE_F32);
    CHECK(d.out_layout == ATR_LAYOUT_XZY && d.bin == 1.0f);

    /* a valid series: the estimate is positive and grows with the volume */
    for (i = 0; i < N; i++) tilts[i] = -50.0f + 10.0f * (float)i;
    d.proj_x = 64;
    d.proj_y = 8;
    d.n_tilts = N;
    d.tilt_deg = tilts;
    d.vol_z = 32;
    d.sart_iters = 2;
    d.sart_subsets = 2;
    CHECK(atr_estimate_device_bytes(&d, &small) == ATR_OK && small > 0);
    d.vol_z = 64;
    CHECK(atr_estimate_device_bytes(&d, &large) == ATR_OK && large > small);

    /* invalid descriptors fail with a message, before any device work */
    d.n_tilts = 0;
    CHECK(atr_estimate_device_bytes(&d, &small) == ATR_ERR_INVALID_ARG);
    CHECK(atr_last_error_string() != NULL && strlen(atr_last_error_string()) > 0);
    {
        atr_plan* plan = NULL;
        CHECK(atr_plan_create(&d, &plan) == ATR_ERR_INVALID_ARG && plan == NULL);
        CHECK(atr_plan_create(NULL, &plan) == ATR_ERR_INVALID_ARG);
    }

    /* MRC: write, read back bit for bit */
    for (i = 0; i < NX * NY * NZ; i++) vol[i] = sinf(0.37f * (float)i) * 100.0f;
    snprintf(path, sizeof(path), "%s/c_consumer.mrc", argv[1]);
    CHECK(atr_mrc_write(path, vol, NX, NY, NZ, 1.5f, 1) == ATR_OK);
    CHECK(atr_mrc_read_info(path, &info) == ATR_OK);
    CHECK(info.nx == NX && info.ny == NY && info.nz == NZ && info.mode == 2);
    CHECK(fabsf(info.pixel_size - 1.5f) < 1e-6f && info.nsymbt == 0 && info.swapped == 0);
    memset(back, 0, sizeof(back));
    CHECK(atr_mrc_read_sections(path, 0, NZ, back) == ATR_OK);
    CHECK(memcmp(vol, back, sizeof(vol)) == 0);
    CHECK(atr_mrc_read_sections(path, NZ - 1, 2, back) != ATR_OK);
    CHECK(strcmp(atr_volume_suffix(0), "_Vol.mrc") == 0 && atr_volume_suffix(5) == NULL);

    /* thickness edges: a slab of high standard deviation in the middle */
    for (i = 0; i < 40; i++) profile[i] = (i >= 12 && i < 28) ? 5.0f : 1.0f;
    CHECK(atr_thickness_edges(profile, 40, 20.0f, &th, NULL) == ATR_OK);
    CHECK(th.low_edge < th.high_edge && th.thickness == th.high_edge - th.low_edge);
    CHECK(th.low_edge > 0 && th.high_edge < 40 * 20);
    CHECK(atr_thickness_edges(profile, 5, 20.0f, &th, NULL) == ATR_ERR_INVALID_ARG);

    printf("ok\n");
    return 0;
}
