/*
 * tws_oracle_body.h -- body of the CPU restatement, included twice by
 * tws_oracle.c with ORACLE_REAL = float and = double.  TEST INFRASTRUCTURE
 * ONLY (see tws_oracle.c).
 */
#define ORACLE_CAT2(a, b) a##_##b
#define ORACLE_CAT(a, b) ORACLE_CAT2(a, b)
#define ORACLE_FN(name) ORACLE_CAT(name, ORACLE_SFX)

typedef ORACLE_REAL ORACLE_CAT(oracle_real, ORACLE_SFX);

/*
 * Radial brush, waterBrush.comp:20-31:
 *   toBrush = BrushPositionTexelCor - gridPosition          (:26, ivec2 -> vec2)
 *   brushDistSq = dot(toBrush, toBrush) / BrushSizeSq        (:27)
 *   a += saturate(1 - brushDistSq) * BrushIntensity          (:28, helper.glsl:68)
 * One invocation per texel of the W x H domain (Terrain.cpp:167 dispatches
 * the whole grid).  The brush parameters are float uniforms, so they enter
 * the double instantiation as the same float values.
 */
static void ORACLE_FN(brush_rect)(int W, int bw, int bh, ORACLE_REAL* d, float cx, float cy,
                                  float intensity, float radius_sq)
{
    const ORACLE_REAL one = (ORACLE_REAL)1;
    const ORACLE_REAL zero = (ORACLE_REAL)0;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < bh; ++y) {
        for (int x = 0; x < bw; ++x) {
            ORACLE_REAL tx = (ORACLE_REAL)cx - (ORACLE_REAL)x;
            ORACLE_REAL ty = (ORACLE_REAL)cy - (ORACLE_REAL)y;
            ORACLE_REAL dist = tx * tx + ty * ty;
            ORACLE_REAL q = dist / (ORACLE_REAL)radius_sq;
            ORACLE_REAL w = one - q;
            w = w < zero ? zero : (w > one ? one : w);
            d[(int64_t)y * W + x] += w * (ORACLE_REAL)intensity;
        }
    }
}

void ORACLE_FN(oracle_brush)(int W, int H, ORACLE_REAL* d, float cx, float cy,
                             float intensity, float radius_sq)
{
    ORACLE_FN(brush_rect)(W, W, H, d, cx, cy, intensity, radius_sq);
}

/*
 * n steps of the virtual-pipe model on a W x H grid.
 *   b    [W*H]    terrain height (TerrainData.r), read only
 *   d    [W*H]    water depth (TerrainData.a), updated in place
 *   f    [W*H*4]  outgoing flux (Flow.xyzw = +x,-x,+y,-y), updated in place
 *   v    [W*H*2]  directed flow of the last step (FlowMap.xy), may be NULL
 *   rain [4]      {cx, cy, intensity, radius_sq}: brush applied before every
 *                 step, may be NULL
 * boundary:
 *   CLOSED  -- a pipe that leaves the domain carries nothing, either way.
 *   ZERO    -- the reference's own edge: a texel outside the image reads as
 *              r=g=b=0 and a=ghost_alpha, both in TerrainData and in Flow
 *              (flowUpdate.comp:18 and flowApply.comp:20 load texel -1 and W
 *              for the edge work groups); ghosts never change.
 *   DISPATCH -- ZERO, plus the reference's dispatch shapes: both passes run
 *              W/16 x H/16 work groups of 16 x 16 texels (Terrain.cpp:258,264)
 *              and the brush W/32 x H/32 groups of 32 x 32 (Terrain.cpp:167),
 *              integer divisions, so texels in columns >= 16*(W/16) or rows
 *              >= 16*(H/16) keep their depth and flux (and are never given a
 *              FlowMap value: v stays as passed in), while their neighbours
 *              read them as ordinary texels; the brush reaches only columns
 *              < 32*(W/32) and rows < 32*(H/32).
 * Returns 0, or -1 if the scratch allocation fails.
 */
int ORACLE_FN(oracle_step)(int W, int H, const ORACLE_REAL* b, ORACLE_REAL* d, ORACLE_REAL* f,
                           ORACLE_REAL F, ORACLE_REAL K, ORACLE_REAL C, int boundary,
                           ORACLE_REAL ghost_alpha, const float* rain, int nsteps,
                           ORACLE_REAL* v)
{
    const int64_t n = (int64_t)W * H;
    const ORACLE_REAL zero = (ORACLE_REAL)0;
    if (n <= 0 || nsteps <= 0) return 0;
    ORACLE_REAL* fn = (ORACLE_REAL*)malloc(sizeof(ORACLE_REAL) * (size_t)(n * 4));
    if (!fn) return -1;
    const int closed = (boundary == TWS_ORACLE_BOUNDARY_CLOSED);
    const int dispatch = (boundary == TWS_ORACLE_BOUNDARY_DISPATCH);
    const int uw = dispatch ? W / 16 * 16 : W, uh = dispatch ? H / 16 * 16 : H;
    const int bw = dispatch ? W / 32 * 32 : W, bh = dispatch ? H / 32 * 32 : H;

    for (int step = 0; step < nsteps; ++step) {
        if (rain) ORACLE_FN(brush_rect)(W, bw, bh, d, rain[0], rain[1], rain[2], rain[3]);

        /* ---- flux pass, flowUpdate.comp:29-62 ---------------------------- */
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y) {
            for (int x = 0; x < W; ++x) {
                const int64_t i = (int64_t)y * W + x;
                if (x >= uw || y >= uh) {  /* no work group covers this texel */
                    for (int c = 0; c < 4; ++c) fn[i * 4 + c] = f[i * 4 + c];
                    continue;
                }
                const int in_x1 = x + 1 < W, in_x0 = x > 0, in_y1 = y + 1 < H, in_y0 = y > 0;
                ORACLE_REAL own = d[i] + b[i];                                   /* :34 */
                ORACLE_REAL hx1 = in_x1 ? d[i + 1] + b[i + 1] : ghost_alpha + zero; /* :37 */
                ORACLE_REAL hx0 = in_x0 ? d[i - 1] + b[i - 1] : ghost_alpha + zero; /* :38 */
                ORACLE_REAL hy1 = in_y1 ? d[i + W] + b[i + W] : ghost_alpha + zero; /* :39 */
                ORACLE_REAL hy0 = in_y0 ? d[i - W] + b[i - W] : ghost_alpha + zero; /* :40 */
                ORACLE_REAL g[4] = {own - hx1, own - hx0, own - hy1, own - hy0}; /* :44-47 */
                ORACLE_REAL o[4];
                for (int c = 0; c < 4; ++c) {
                    ORACLE_REAL t = f[i * 4 + c] * F + g[c] * K;                 /* :53 */
                    o[c] = ORACLE_MAX(zero, t);                                  /* :54 */
                }
                if (closed) {
                    if (!in_x1) o[0] = zero;
                    if (!in_x0) o[1] = zero;
                    if (!in_y1) o[2] = zero;
                    if (!in_y0) o[3] = zero;
                }
                /* :57 "x + y + z + w": Mesa (the executable reference) sums
                 * the four components pairwise; the llvmpipe fixtures pin
                 * this order (tests/test_golden.py). */
                ORACLE_REAL total = ((o[0] + o[1]) + (o[2] + o[3])) * C;         /* :57 */
                if (total > d[i]) {                                              /* :58 */
                    ORACLE_REAL s = d[i] / total;                                /* :59 */
                    for (int c = 0; c < 4; ++c) o[c] = o[c] * s;
                }
                for (int c = 0; c < 4; ++c) fn[i * 4 + c] = o[c];                /* :62 */
            }
        }

        /* ---- apply pass, flowApply.comp:32-52 ----------------------------- */
        const int last = (step == nsteps - 1);
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y) {
            for (int x = 0; x < W; ++x) {
                const int64_t i = (int64_t)y * W + x;
                if (x >= uw || y >= uh) continue;  /* no work group covers this texel */
                const ORACLE_REAL* o = fn + i * 4;
                ORACLE_REAL gz = closed ? zero : ghost_alpha;  /* ghost Flow.w */
                ORACLE_REAL fx1 = x + 1 < W ? fn[(i + 1) * 4 + 1] : zero;        /* :32 X1.y */
                ORACLE_REAL fx0 = x > 0 ? fn[(i - 1) * 4 + 0] : zero;            /* :33 X0.x */
                ORACLE_REAL fy1 = y + 1 < H ? fn[(i + W) * 4 + 3] : gz;          /* :34 Y1.w */
                ORACLE_REAL fy0 = y > 0 ? fn[(i - W) * 4 + 2] : zero;            /* :35 Y0.z */
                ORACLE_REAL in = ((fx1 + fx0) + fy1) + fy0;                      /* :38 */
                ORACLE_REAL out = ((o[0] + o[1]) + o[2]) + o[3];                 /* :39 */
                ORACLE_REAL nd = d[i] + (in - out) * C;                          /* :41 */
                d[i] = ORACLE_MAX(zero, nd);
                if (last && v) {
                    v[i * 2 + 0] = (fx1 - o[0]) - (fx0 - o[1]);                  /* :45 */
                    v[i * 2 + 1] = (fy1 - o[2]) - (fy0 - o[3]);                  /* :46 */
                }
            }
        }
        memcpy(f, fn, sizeof(ORACLE_REAL) * (size_t)(n * 4));
    }
    free(fn);
    return 0;
}

#undef ORACLE_FN
#undef ORACLE_CAT
#undef ORACLE_CAT2
