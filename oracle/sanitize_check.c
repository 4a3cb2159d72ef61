/*
 * Sanitizer driver for the CPU oracle (test infrastructure): built with
 * -fsanitize=address,undefined by `make -C oracle sanitize` and run by
 * tests/test_sanitizers.py.  It renders small scenes through oracle_render
 * (several threads, uneven resolutions, every output and diagnostic buffer)
 * and checks closed forms, so that any out-of-bounds access, use of
 * uninitialized memory that ASan tracks, or undefined arithmetic in the oracle
 * aborts the run.
 *
 * Closed form checked: a single isotropic Gaussian straight ahead of a camera
 * composites at its centre pixel with alpha = min(0.99, opacity) and depth = z.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vgr_oracle.h"

static int failures = 0;

static void expect(int condition, const char* what) {
  if (!condition) {
    fprintf(stderr, "sanitize_check: FAILED %s\n", what);
    ++failures;
  }
}

static void identity4(float* m) {
  memset(m, 0, 16 * sizeof(float));
  m[0] = m[5] = m[10] = m[15] = 1.0f;
}

static void pinhole(float* k, float focal, float cx, float cy) {
  memset(k, 0, 9 * sizeof(float));
  k[0] = focal;
  k[4] = focal;
  k[2] = cx;
  k[5] = cy;
  k[8] = 1.0f;
}

/* Single Gaussian on the optical axis at depth 4 (64x48 camera at the origin). */
static void single_gaussian_kat(void) {
  const float means[3] = {0.0f, 0.0f, 4.0f};
  const float scales[3] = {0.05f, 0.05f, 0.05f};
  const float rotations[4] = {1.0f, 0.0f, 0.0f, 0.0f};
  const float opacity = 0.8f;
  const float features[3] = {0.25f, 0.5f, 0.75f};
  const int64_t semantic = 17, registered = 0, offsets[2] = {0, 1}, camera_scene = 0;
  float covariances[6];
  expect(oracle_precompute_covariances(scales, rotations, 1, covariances) == 0, "covariance precompute");
  float view[16], env[16], k[9];
  identity4(view);
  identity4(env);
  pinhole(k, 60.0f, 32.0f, 24.0f);
  oracle_scene scene = {means, covariances, &opacity, features, &semantic, &registered, offsets, 1, 3, 1, 0, 0};
  oracle_cameras cameras = {view, k, env, &camera_scene, NULL, 1, 1, NULL, NULL, NULL};
  oracle_options options = {48, 64, 0.01f, 100.0f, 3.33f, 0.3f, 0.01f, 0, 16, 0, 1, 2};
  const int pixels = 64 * 48;
  float* rgb = calloc((size_t)pixels * 3, sizeof(float));
  float* depth = calloc((size_t)pixels, sizeof(float));
  float* alpha = calloc((size_t)pixels, sizeof(float));
  int64_t* labels = calloc((size_t)pixels, sizeof(int64_t));
  int32_t tile_counts[12];
  uint64_t keys[64];
  int32_t sorted[64];
  oracle_outputs outputs;
  memset(&outputs, 0, sizeof(outputs));
  outputs.rgb = rgb;
  outputs.depth = depth;
  outputs.alpha = alpha;
  outputs.semantic = labels;
  outputs.tile_counts = tile_counts;
  outputs.sorted_keys = keys;
  outputs.sorted_gaussians = sorted;
  outputs.sorted_capacity = 64;
  outputs.global_tiles = 12;
  expect(oracle_render(&scene, &cameras, &options, &outputs) == 0, "single-Gaussian render");
  /* Pixel (32, 24) has its centre at (32.5, 24.5): take the pixel nearest the mean (32, 24). */
  const int centre = 24 * 64 + 31;
  expect(outputs.counters[0] == 1, "one visible record");
  expect(fabsf(depth[centre] - 4.0f) < 1e-5f, "centre depth = z");
  expect(alpha[centre] > 0.5f && alpha[centre] <= 0.8f, "centre alpha near the opacity");
  expect(labels[centre] == 17, "centre semantic id");
  free(rgb);
  free(depth);
  free(alpha);
  free(labels);
}

/* Random scene, two scenes, uneven resolutions, several threads. */
static void random_scene_render(void) {
  const int count = 3000, batch = 3;
  float* means = malloc(sizeof(float) * 3 * count);
  float* scales = malloc(sizeof(float) * 3 * count);
  float* rotations = malloc(sizeof(float) * 4 * count);
  float* opacities = malloc(sizeof(float) * count);
  float* features = malloc(sizeof(float) * 12 * count);
  int64_t* semantic = malloc(sizeof(int64_t) * count);
  float* covariances = malloc(sizeof(float) * 6 * count);
  unsigned state = 12345u;
#define NEXT() ((state = state * 1664525u + 1013904223u), (float)(state >> 8) / 16777216.0f)
  for (int g = 0; g < count; ++g) {
    for (int i = 0; i < 3; ++i) means[3 * g + i] = 6.0f * (NEXT() - 0.5f);
    means[3 * g + 2] += 8.0f;
    for (int i = 0; i < 3; ++i) scales[3 * g + i] = 0.01f + 0.05f * NEXT();
    for (int i = 0; i < 4; ++i) rotations[4 * g + i] = NEXT() - 0.5f;
    opacities[g] = NEXT();
    for (int i = 0; i < 12; ++i) features[12 * g + i] = NEXT() - 0.3f;
    semantic[g] = g % 50;
  }
#undef NEXT
  expect(oracle_precompute_covariances(scales, rotations, count, covariances) == 0, "random covariances");
  const int64_t registered[2] = {3, 9}, offsets[3] = {0, 1800, count}, camera_scene[3] = {3, 9, 3};
  float views[48], envs[48], ks[27];
  const int32_t widths[3] = {40, 33, 17}, heights[3] = {24, 31, 9};
  int64_t pixel_offsets[3];
  int64_t pixels = 0;
  for (int b = 0; b < batch; ++b) {
    identity4(views + 16 * b);
    identity4(envs + 16 * b);
    views[16 * b + 3] = 0.3f * (float)b;
    pinhole(ks + 9 * b, 30.0f + 5.0f * (float)b, 0.5f * (float)widths[b], 0.5f * (float)heights[b]);
    pixel_offsets[b] = pixels;
    pixels += (int64_t)widths[b] * heights[b];
  }
  oracle_scene scene = {means, covariances, opacities, features, semantic, registered, offsets, count, 12, 2, 1, 0};
  oracle_cameras cameras = {views, ks, envs, camera_scene, NULL, batch, batch, widths, heights, pixel_offsets};
  oracle_options options = {1, 1, 0.01f, 100.0f, 3.33f, 0.3f, 0.01f, 1, 8, 1, 1, 3};
  float* rgb = calloc((size_t)pixels * 3, sizeof(float));
  float* depth = calloc((size_t)pixels, sizeof(float));
  float* alpha = calloc((size_t)pixels, sizeof(float));
  int64_t* labels = calloc((size_t)pixels, sizeof(int64_t));
  oracle_outputs outputs;
  memset(&outputs, 0, sizeof(outputs));
  outputs.rgb = rgb;
  outputs.depth = depth;
  outputs.alpha = alpha;
  outputs.semantic = labels;
  expect(oracle_render(&scene, &cameras, &options, &outputs) == 0, "random-scene render");
  expect(outputs.counters[0] > 0 && outputs.counters[1] >= outputs.counters[0], "counters plausible");
  int finite = 1;
  for (int64_t i = 0; i < pixels; ++i) {
    finite &= isfinite(rgb[3 * i]) && alpha[i] >= 0.0f && alpha[i] <= 1.0f;
  }
  expect(finite, "finite rgb, alpha in [0, 1]");
  free(rgb);
  free(depth);
  free(alpha);
  free(labels);
  free(means);
  free(scales);
  free(rotations);
  free(opacities);
  free(features);
  free(semantic);
  free(covariances);
}

int main(void) {
  single_gaussian_kat();
  random_scene_render();
  if (failures != 0) return 1;
  printf("sanitize_check: oracle OK\n");
  return 0;
}
