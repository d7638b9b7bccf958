/*
 * hosek_probe.c -- test-side wrapper over the Hosek-Wilkie 2013 sky model that rtcamp8 vendors
 * (ext/hosek_wilkie_2013, 3-clause BSD).  Built only in the development container, against the
 * vendored sources where they lie, into oracle/_ref/ (see oracle/ref/Makefile).  It provides the
 * FP64 ground truth the FP32 sky bake is checked against, and the per-band model states the bake
 * consumes (renderer/renderer_scene.cpp:570-592 computes the same states on the host).
 */
#include "ArHosekSkyModel.h"

#include <string.h>

extern double* solarDatasets[];
extern double* limbDarkeningDatasets[];

/* Flattened ArHosekSkyModelState: configs[11][9], radiances[11], turbidity, solar_radius,
 * emission_correction_factor_sky[11], emission_correction_factor_sun[11], albedo, elevation. */
#define HOSEK_STATE_DOUBLES (99 + 11 + 2 + 11 + 11 + 2)

int hosek_state_size(void) { return HOSEK_STATE_DOUBLES; }

void hosek_state(double elevation, double turbidity, double albedo, double* out)
{
    ArHosekSkyModelState* s = arhosekskymodelstate_alloc_init(elevation, turbidity, albedo);
    int k = 0;
    for (int i = 0; i < 11; ++i)
        for (int j = 0; j < 9; ++j) out[k++] = s->configs[i][j];
    for (int i = 0; i < 11; ++i) out[k++] = s->radiances[i];
    out[k++] = s->turbidity;
    out[k++] = s->solar_radius;
    for (int i = 0; i < 11; ++i) out[k++] = s->emission_correction_factor_sky[i];
    for (int i = 0; i < 11; ++i) out[k++] = s->emission_correction_factor_sun[i];
    out[k++] = s->albedo;
    out[k++] = s->elevation;
    arhosekskymodelstate_free(s);
}

double hosek_radiance(double elevation, double turbidity, double albedo, double theta, double gamma, double wl)
{
    ArHosekSkyModelState* s = arhosekskymodelstate_alloc_init(elevation, turbidity, albedo);
    double v = arhosekskymodel_radiance(s, theta, gamma, wl);
    arhosekskymodelstate_free(s);
    return v;
}

double hosek_solar_radiance(double elevation, double turbidity, double albedo, double theta, double gamma, double wl)
{
    ArHosekSkyModelState* s = arhosekskymodelstate_alloc_init(elevation, turbidity, albedo);
    double v = arhosekskymodel_solar_radiance(s, theta, gamma, wl);
    arhosekskymodelstate_free(s);
    return v;
}

/* One wavelength channel of the solar radiance table (1800 = 10 turbidities x 45 pieces x 4). */
void hosek_solar_table(int wl, double* out)
{
    memcpy(out, solarDatasets[wl], sizeof(double) * 1800);
}

void hosek_limb_table(int wl, double* out)
{
    memcpy(out, limbDarkeningDatasets[wl], sizeof(double) * 6);
}

extern double* datasets[];
extern double* datasetsRad[];

/* One wavelength channel of the spectral model's configuration table (1080 = 2 albedos x 10 turbidities
 * x 6 elevation control points x 9 parameters) and of its radiance table (120 = 2 x 10 x 6). */
void hosek_config_table(int wl, double* out)
{
    memcpy(out, datasets[wl], sizeof(double) * 1080);
}

void hosek_radiance_table(int wl, double* out)
{
    memcpy(out, datasetsRad[wl], sizeof(double) * 120);
}
