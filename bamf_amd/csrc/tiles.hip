// Frame tiling for the chain batch and placement of each tile's posterior
// grid in the stitched frame image (@RJ/makeSubRegions.m:42-102,
// @RJ/removeOverlapping.m:29-54, @RJ/stitchROIs.m:28-65), on the device.
//
// Geometry (1-based, as the MATLAB): Xsub = ceil(XPix / Box) tiles across,
// Ysub = ceil(YPix / Box) down.  Tile ix spans columns sx .. ex with
// sx = Box (ix-1) - (Overlap-1), ex = Box ix + Overlap, clamped to 1 and
// Box + Overlap for the first tile and to XPix for the last; its width is
// ex - sx + 1 (XPix when Xsub = 1).  Rows likewise.  Tiles are numbered
// frame-major, then column-major over positions: Ind = Xsub Ysub (frame-1) +
// Ysub (ix-1) + iy (makeSubRegions.m:90).
//
// The chain batch stores tile b compactly (column-major, x * ysize + y) at
// pix_offset[b]; the offsets have a closed form because a tile's size depends
// only on (ix, iy): within a frame, tile (ix, iy) starts after the columns of
// tiles 1..ix-1 (PW(ix) columns of SH rows each, SH = sum of the tile heights)
// and the rows of tiles (ix, 1..iy-1) of its own width.
//
// Placement in the stitched image (stitchROIs.m with ImSize(3) = 1, as
// RJ.m:233-234 calls it): grid pixel (y', x') of tile (ix, iy) lands at image
// row Zoom (sy-1) + y', column Zoom (sx-1) + x' of a (Zoom YPix) x (Zoom XPix)
// column-major image; removeOverlapping zeroes Zoom*Overlap grid columns at
// each interior side, so only the core window of each tile is kept.  The
// cores partition the image.
#include "common.hpp"

namespace bamf {
namespace {

struct TileGeo {
    int ypix, xpix, nframes, box, ov, xsub, ysub;
    int64_t sw, sh;  // sum of tile widths across, of tile heights down
};

// 1-based tile start and size along one axis (makeSubRegions.m:62-87)
__host__ __device__ inline int tile_start(int i, int box, int ov) { return i == 1 ? 1 : box * (i - 1) - (ov - 1); }
__host__ __device__ inline int tile_size(int i, int nsub, int pix, int box, int ov) {
    if (nsub == 1) return pix;
    if (i == 1) return box + ov;
    if (i == nsub) return pix - box * (nsub - 1) + ov;
    return box + 2 * ov;
}
// sum of the sizes of tiles 1 .. i-1 (none of them the last)
__host__ __device__ inline int64_t tile_prefix(int i, int box, int ov) {
    return i <= 1 ? 0 : (int64_t)(box + ov) + (int64_t)(i - 2) * (box + 2 * ov);
}

// One work-group per tile column (frame f, position ix: the Ysub tiles that
// share a column range, contiguous in the compact batch, tile iy's W x H(iy)
// block after tile iy-1's): metadata by the first Ysub threads, then the
// pixels, one stacked row per work-item, several loads in flight per thread.
constexpr int kTileThreads = 256, kTileUnroll = 8;
__global__ __launch_bounds__(kTileThreads) void k_tiles(const float* __restrict__ data, TileGeo g,
                                                        float* __restrict__ pixels, int64_t* __restrict__ pix_offset,
                                                        int32_t* __restrict__ xsize, int32_t* __restrict__ ysize,
                                                        int32_t* __restrict__ frame, int32_t* __restrict__ roi_ind,
                                                        float* __restrict__ roi_data, int zoom,
                                                        int64_t* __restrict__ grid_offset,
                                                        int32_t* __restrict__ grid_place) {
    const int f = blockIdx.x / g.xsub;
    const int ix = blockIdx.x - f * g.xsub + 1;
    const int w = tile_size(ix, g.xsub, g.xpix, g.box, g.ov), sx = tile_start(ix, g.box, g.ov);
    const int64_t col0 = (int64_t)f * g.sw * g.sh + tile_prefix(ix, g.box, g.ov) * g.sh;  // compact offset
    const int64_t b0 = ((int64_t)f * g.xsub + (ix - 1)) * g.ysub;                          // first tile (0-based)
    for (int iy = threadIdx.x + 1; iy <= g.ysub; iy += blockDim.x) {
        const int64_t b = b0 + iy - 1;
        const int h = tile_size(iy, g.ysub, g.ypix, g.box, g.ov), sy = tile_start(iy, g.box, g.ov);
        if (pix_offset) pix_offset[b] = col0 + (int64_t)w * tile_prefix(iy, g.box, g.ov);
        if (xsize) xsize[b] = w;
        if (ysize) ysize[b] = h;
        if (frame) frame[b] = f + 1;
        if (roi_ind) roi_ind[b] = g.ysub * (ix - 1) + iy;
        if (grid_offset) {
            const int64_t ld = (int64_t)zoom * g.ypix;
            grid_offset[b] = (int64_t)zoom * (sx - 1) * ld + (int64_t)zoom * (sy - 1);
            int32_t* gp = grid_place + 5 * b;
            const int zo = zoom * g.ov, wz = zoom * w, hz = zoom * h;
            gp[0] = (int32_t)ld;
            gp[1] = ix != 1 ? zo : 0;
            gp[2] = ix != g.xsub ? wz - zo : wz;
            gp[3] = iy != 1 ? zo : 0;
            gp[4] = iy != g.ysub ? hz - zo : hz;
        }
    }
    const float* src = data + ((int64_t)f * g.xpix + (sx - 1)) * g.ypix;  // column sx of frame f, row 1
    if (pixels) {
        // Work items (stacked row R, chunk of kTileUnroll columns): R runs
        // over the tiles' rows one after another (tiles 2 .. Ysub-1 all have
        // height hm, so R's tile has a closed form); consecutive work-items
        // take consecutive rows of one chunk, so the frame loads are coalesced
        // across rows and each work-item keeps kTileUnroll loads in flight;
        // the compact writes are contiguous within each tile's column.
        const int h1 = tile_size(1, g.ysub, g.ypix, g.box, g.ov), hm = g.box + 2 * g.ov;
        const int sh = (int)g.sh, nch = (w + kTileUnroll - 1) / kTileUnroll;
        for (int item = threadIdx.x; item < sh * nch; item += kTileThreads) {
            const int c = item / sh, R = item - c * sh;
            int iy, y;
            if (R < h1 || g.ysub == 1) {
                iy = 1; y = R;
            } else {
                iy = min((R - h1) / hm + 2, g.ysub);
                y = R - h1 - (iy - 2) * hm;
            }
            const int h = tile_size(iy, g.ysub, g.ypix, g.box, g.ov);
            const int x0 = c * kTileUnroll, nx = min(kTileUnroll, w - x0);
            const float* sp = src + (int64_t)x0 * g.ypix + tile_start(iy, g.box, g.ov) - 1 + y;
            float* dp = pixels + col0 + (int64_t)w * tile_prefix(iy, g.box, g.ov) + (int64_t)x0 * h + y;
            if (nx == kTileUnroll) {
                float v[kTileUnroll];
#pragma unroll
                for (int u = 0; u < kTileUnroll; u++) v[u] = sp[(int64_t)u * g.ypix];
#pragma unroll
                for (int u = 0; u < kTileUnroll; u++) dp[(int64_t)u * h] = v[u];
            } else {
                for (int u = 0; u < nx; u++) dp[(int64_t)u * h] = sp[(int64_t)u * g.ypix];
            }
        }
    }
    if (roi_data) {  // ROI_Data(1:ROIey, 1:ROIex, Ind), zero elsewhere (makeSubRegions.m:45,95)
        const int side = g.box + 2 * g.ov;
        const int64_t ne = (int64_t)g.ysub * side * side;
        float* dst = roi_data + b0 * side * side;
        for (int64_t e = threadIdx.x; e < ne; e += kTileThreads) {
            const int t = (int)(e / (side * side));
            const int r = (int)(e - (int64_t)t * side * side);
            const int x = r / side, y = r - x * side;
            const int iy = t + 1, h = tile_size(iy, g.ysub, g.ypix, g.box, g.ov);
            dst[e] = (x < w && y < h) ? src[(int64_t)x * g.ypix + tile_start(iy, g.box, g.ov) - 1 + y] : 0.0f;
        }
    }
}

bool geo_of(int ypix, int xpix, int nframes, int box, int overlap, TileGeo* g) {
    if (ypix <= 0 || xpix <= 0 || nframes < 0 || box <= 0 || overlap < 0) return false;
    g->ypix = ypix; g->xpix = xpix; g->nframes = nframes; g->box = box; g->ov = overlap;
    g->xsub = (xpix + box - 1) / box;
    g->ysub = (ypix + box - 1) / box;
    g->sw = g->sh = 0;
    for (int i = 1; i <= g->xsub; i++) {
        const int w = tile_size(i, g->xsub, xpix, box, overlap), s = tile_start(i, box, overlap);
        if (w <= 0 || s < 1 || s + w - 1 > xpix) return false;  // a tile outside the frame
        g->sw += w;
    }
    for (int i = 1; i <= g->ysub; i++) {
        const int h = tile_size(i, g->ysub, ypix, box, overlap), s = tile_start(i, box, overlap);
        if (h <= 0 || s < 1 || s + h - 1 > ypix) return false;
        g->sh += h;
    }
    return true;
}

}  // namespace
}  // namespace bamf

using namespace bamf;

extern "C" {

int64_t bamf_subregion_count(int ypix, int xpix, int nframes, int box, int overlap) {
    TileGeo g;
    if (!geo_of(ypix, xpix, nframes, box, overlap, &g)) return -1;
    return (int64_t)g.xsub * g.ysub * nframes;
}

int64_t bamf_subregion_pixels(int ypix, int xpix, int nframes, int box, int overlap) {
    TileGeo g;
    if (!geo_of(ypix, xpix, nframes, box, overlap, &g)) return -1;
    return g.sw * g.sh * nframes;
}

int bamf_make_subregions(const float* d_data, int ypix, int xpix, int nframes, int box, int overlap,
                         float* d_pixels, int64_t* d_pix_offset, int32_t* d_xsize, int32_t* d_ysize,
                         int32_t* d_frame, int32_t* d_roi_ind, float* d_roi_data, int zoom,
                         int64_t* d_grid_offset, int32_t* d_grid_place, void* stream) {
    TileGeo g;
    BAMF_CHECK_ARG(geo_of(ypix, xpix, nframes, box, overlap, &g),
                   "bad tiling (%d x %d frames, box %d, overlap %d): every tile must lie inside the frame", ypix,
                   xpix, box, overlap);
    BAMF_CHECK_ARG(!(d_grid_offset || d_grid_place) || (d_grid_offset && d_grid_place && zoom >= 1),
                   "grid placement needs both arrays and zoom >= 1");
    BAMF_CHECK_ARG(!d_grid_offset || (int64_t)zoom * ypix <= 0x7fffffff, "zoomed frame too tall");
    const int64_t n = (int64_t)g.xsub * g.ysub * nframes;
    if (n == 0) return BAMF_OK;
    BAMF_CHECK_ARG(d_data || !(d_pixels || d_roi_data), "null frames");
    BAMF_CHECK_ARG((int64_t)g.xsub * nframes <= 0x7fffffff, "too many tile columns");
    BAMF_CHECK_ARG((int64_t)g.sh * (box + 2 * overlap) <= 0x7fffffff, "frames too tall for the tile copy");
    hipLaunchKernelGGL(k_tiles, dim3((unsigned)((int64_t)g.xsub * nframes)), dim3(kTileThreads), 0,
                       as_stream(stream), d_data, g, d_pixels,
                       d_pix_offset, d_xsize, d_ysize, d_frame, d_roi_ind, d_roi_data, zoom, d_grid_offset,
                       d_grid_place);
    BAMF_LAUNCH_CHECK("k_tiles");
    return BAMF_OK;
}

}  // extern "C"
