// mbr_nogui.cpp — command-line renderer over the C ABI, the -nogui driver of
// the reference (ReadComandlineParams, src/interface.cpp:3334-3561; the frame
// loop of MainRender, src/Render3D.cpp:1995-2470): load a settings file with
// overrides, render one frame or an animation (flight path or keyframes),
// compile each image, apply the post effects, write it.
//
//   mbr_nogui [options...] settings_file
//     -nogui              accepted (this driver has no GUI)
//     -o key=value        override a setting of the file
//     -res WIDTHxHEIGHT   override the image resolution
//     -fpk N              override frames_per_keyframe
//     -format FORMAT      jpg | png | png16 | png16alpha
//     -start N, -end N    frame range (animation modes)
//     -flight             play the flight path file_animation_path (one line
//                         "vp_x vp_y vp_z alfa beta gamma" per frame, radians)
//     -keyframe           interpolate the keyframes file_keyframes%05d.fract
//                         (Catmull-Rom, frames_per_keyframe per key)
//     -precision P        fp32 (default) | fp64 | fp32march   (this driver)
//     -out PATH           write the image to PATH instead of
//                         file_destination + 00000 + extension (this driver)
//     -batch N            animation frames per batched launch (default 16;
//                         1 renders frame by frame)              (this driver)
//
// The image goes to file_destination + frame number (%05d) + ".png", as
// IndexFilename (src/files.cpp:30-35) names it; with the settings' tiles > 1
// the frame is rendered tile by tile and joined into file_destination +
// "_fromTiles.png" (16-bit).  A frame or tile whose output file exists is
// skipped, as in the reference's -nogui mode, so a second run resumes an
// interrupted animation or tiled frame.  There is no JPEG encoder in
// this build: -format jpg writes an 8-bit PNG.  Network rendering (-ip,
// -port) is not part of the render core and is rejected.
//
// Animations render in batches: up to -batch consecutive frames still to do
// go to mbr_render_batch as one HIP graph launch into device framebuffers,
// and each frame is then compiled and written as a single render's would be
// (the results equal frame-by-frame rendering).  Scenes the batch does not
// take (aux lights, volumetric lights, iteration fog, reflections) and tiled
// frames render frame by frame.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "mbr.h"

namespace {

void usage() {
  std::printf(
      "Syntax:\nmbr_nogui [options...] settings_file\noptions:\n"
      "  -nogui            - accepted; this driver never opens a GUI\n"
      "  -o key=value      - override item 'key' from settings file with new value 'value'\n"
      "  -res WIDTHxHEIGHT - override image resolution\n"
      "  -fpk N            - override frames per key parameter\n"
      "  -start N          - start rendering from frame number N\n"
      "  -end N            - rendering will end on frame number N\n"
      "  -flight           - render the flight path (file_animation_path)\n"
      "  -keyframe         - render the keyframe animation (file_keyframes)\n"
      "  -format FORMAT    - image output format: jpg (written as 8-bit PNG), png, png16, png16alpha\n"
      "  -precision P      - fp32 (FP64 march, FP32 shading; default), fp64, fp32march\n"
      "  -out PATH         - output file (default: file_destination + frame + .png)\n"
      "  -batch N          - animation frames per batched GPU launch (default 16, 1 = frame by frame)\n");
}

void make_parent(const std::string &path) {
  for (size_t i = 1; i < path.size(); i++) {
    if (path[i] == '/') mkdir(path.substr(0, i).c_str(), 0755);
  }
}

// IndexFilename (src/files.cpp:30-35)
std::string index_filename(const char *prefix, const char *ext, int number) {
  char idx[16];
  std::snprintf(idx, sizeof idx, "%05d", number);
  return std::string(prefix) + idx + "." + ext;
}

bool file_exists(const std::string &path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0;
}

// The host planes of one rendered frame.
struct Planes {
  int w = 0, h = 0;
  std::vector<float> r, g, b, z;
  std::vector<uint8_t> cr, cg, cb;
  std::vector<uint16_t> alpha, opacity;
  void resize(int width, int height) {
    w = width;
    h = height;
    const size_t n = (size_t)w * h;
    r.resize(n); g.resize(n); b.resize(n); z.resize(n);
    cr.resize(n); cg.resize(n); cb.resize(n);
    alpha.resize(n); opacity.resize(n);
  }
};

// the reference's end-of-render report (src/Render3D.cpp:638-652)
void report(const mbr_stats &st, double sec, size_t rays) {
  std::printf("rendering time: %.3f s (GPU %.3f ms), %.2f Mrays/s\n", sec, st.ms_total, rays / sec / 1e6);
  std::printf("hits %llu of %llu rays, DE calls %llu, iterations %llu (%.1f per DE call), missed DE %.4f%%\n",
              (unsigned long long)st.hits, (unsigned long long)st.rays,
              (unsigned long long)(st.march_de_calls + st.shade_de_calls), (unsigned long long)st.iterations,
              (double)st.iterations / (double)std::max<uint64_t>(1, st.march_de_calls + st.shade_de_calls),
              100.0 * (double)st.missed_de / (double)std::max<uint64_t>(1, st.hits));
}

// Compile a rendered frame, apply the post effects and write it to `out`
// (or file_destination + index).  With tile_out set the compiled 16-bit
// image is handed back instead of written (the tile loop stores it).
// 0 on success.
int finish_image(mbr_ctx *ctx, const mbr_scene &s, Planes &pl, const std::string &format, std::string out, int index,
                 std::vector<uint16_t> *tile_out) {
  const int w = pl.w, h = pl.h;
  const size_t n = (size_t)w * h;
  std::vector<uint16_t> img16(3 * n);
  const float *planes[3] = {pl.r.data(), pl.g.data(), pl.b.data()};
  int rc = mbr_compile_image(&s, w, h, planes, img16.data());
  // post effects in the reference's order (src/Render3D.cpp:1094-1105)
  if (rc == MBR_OK && s.ssao_enabled) {
    mbr_post_planes pp;
    std::memset(&pp, 0, sizeof pp);
    pp.width = w;
    pp.height = h;
    pp.z = pl.z.data();
    pp.opacity = pl.opacity.data();
    pp.colour[0] = pl.cr.data();
    pp.colour[1] = pl.cg.data();
    pp.colour[2] = pl.cb.data();
    pp.image16 = img16.data();
    pp.alpha = pl.alpha.data();
    rc = mbr_post_ssao(ctx, &s, &pp, 0, nullptr);
    if (rc != MBR_OK) std::printf("SSAO failed: %s\n", mbr_last_error(ctx));
    else std::printf("SSAO rendered (quality %d)\n", s.ssao_quality);
  }
  if (rc == MBR_OK && s.dof_enabled) {
    mbr_post_planes pp;
    std::memset(&pp, 0, sizeof pp);
    pp.width = w;
    pp.height = h;
    pp.z = pl.z.data();
    pp.image16 = img16.data();
    pp.alpha = pl.alpha.data();
    rc = mbr_post_dof(ctx, &s, &pp, 0, nullptr);
    if (rc != MBR_OK) std::printf("DOF failed: %s\n", mbr_last_error(ctx));
    else std::printf("DOF rendered (focus %g, radius %g)\n", s.dof_focus, s.dof_radius);
  }
  if (tile_out) {
    if (rc == MBR_OK) tile_out->swap(img16);
    return rc == MBR_OK ? 0 : 1;
  }
  if (out.empty()) out = index_filename(s.file_destination, "png", index);
  make_parent(out);
  const int bits = (format == "png16" || format == "png16alpha") ? 16 : 8;
  if (rc == MBR_OK)
    rc = mbr_write_png(out.c_str(), w, h, img16.data(), format == "png16alpha" ? pl.alpha.data() : nullptr, bits);
  if (rc != MBR_OK) {
    std::printf("cannot write %s (status %d)\n", out.c_str(), rc);
    return 1;
  }
  std::printf("image saved: %s%s\n", out.c_str(), format == "jpg" ? " (8-bit PNG: no JPEG encoder in this build)" : "");
  return 0;
}

// Render one image of scene s and finish it (finish_image).  0 on success.
int render_image(mbr_ctx *ctx, const mbr_scene &s, int precision, const std::string &format, std::string out,
                 int index, std::vector<uint16_t> *tile_out) {
  int rc = mbr_set_scene(ctx, &s, nullptr);
  if (rc != MBR_OK) {
    std::printf("scene rejected: %s\n", mbr_last_error(ctx));
    return 1;
  }
  Planes pl;
  pl.resize(s.image_width, s.image_height);
  mbr_framebuffer fb;
  fb.width = pl.w;
  fb.height = pl.h;
  fb.rgb[0] = pl.r.data();
  fb.rgb[1] = pl.g.data();
  fb.rgb[2] = pl.b.data();
  fb.z = pl.z.data();
  fb.colour[0] = pl.cr.data();
  fb.colour[1] = pl.cg.data();
  fb.colour[2] = pl.cb.data();
  fb.alpha = pl.alpha.data();
  fb.opacity = pl.opacity.data();
  mbr_render_opts opts;
  std::memset(&opts, 0, sizeof opts);
  opts.precision = precision;
  mbr_stats st;
  const auto t0 = std::chrono::steady_clock::now();
  rc = mbr_render(ctx, &opts, &fb, &st);
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc != MBR_OK) {
    std::printf("render failed: %s\n", mbr_last_error(ctx));
    return 1;
  }
  report(st, sec, (size_t)pl.w * pl.h);
  return finish_image(ctx, s, pl, format, out, index, tile_out);
}

// One frame (src/Render3D.cpp:2314-2496).  With tiles > 1 the frame is
// rendered as tiles x tiles images of the frame's size, each a tile of the
// full view; each tile's 16-bit image goes to file_destination + "tile" +
// index (raw sRGB16, as the reference saves them), and when the last tile is
// done the tiles are joined into file_destination + "_fromTiles.png", a
// 16-bit PNG of (width x tiles) x (height x tiles), and removed
// (SaveFromTilesPNG16, src/files.cpp:269-373).
//
// Resume: -nogui always saves, and the reference then skips every frame or
// tile whose output file already exists (src/Render3D.cpp:2335-2350), so an
// interrupted animation or tiled render continues where it stopped when run
// again; tiles left by the interrupted run are joined with the new ones.

int render_frame(mbr_ctx *ctx, const mbr_scene &s, int precision, const std::string &format, std::string out,
                 int index) {
  const int tiles = s.tiles < 1 ? 1 : (s.tiles > 100 ? 100 : s.tiles);
  if (tiles == 1) {
    const std::string path = out.empty() ? index_filename(s.file_destination, "png", index) : out;
    if (file_exists(path)) {
      std::printf("Output file (%s) exists. Skip to next file index\n", path.c_str());
      return 0;
    }
    return render_image(ctx, s, precision, format, out, index, nullptr);
  }
  const int w = s.image_width, h = s.image_height;
  for (int tile = 0; tile < tiles * tiles; tile++) {
    const int index2 = index * tiles * tiles + tile;
    if (file_exists(index_filename(s.file_destination, "tile", index2))) {
      std::printf("Output file (%s) exists. Skip to next file index\n",
                  index_filename(s.file_destination, "tile", index2).c_str());
      continue;
    }
    std::printf("---------- Tile: %d -------------\n", index2);
    mbr_scene ts = s;
    ts.tiles = tiles;
    ts.tile_index = tile;
    std::vector<uint16_t> img;
    if (render_image(ctx, ts, precision, format, "", index2, &img) != 0) return 1;
    const std::string path = index_filename(s.file_destination, "tile", index2);
    make_parent(path);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(img.data(), sizeof(uint16_t), img.size(), f) != img.size()) {
      if (f) std::fclose(f);
      std::printf("cannot write %s\n", path.c_str());
      return 1;
    }
    std::fclose(f);
  }
  // join the tiles, row of tiles by row of tiles
  const int W = w * tiles, H = h * tiles;
  std::vector<uint16_t> full((size_t)3 * W * H);
  for (int tile = 0; tile < tiles * tiles; tile++) {
    const std::string path = index_filename(s.file_destination, "tile", index * tiles * tiles + tile);
    FILE *f = std::fopen(path.c_str(), "rb");
    std::vector<uint16_t> img((size_t)3 * w * h);
    const bool ok = f && std::fread(img.data(), sizeof(uint16_t), img.size(), f) == img.size();
    if (f) std::fclose(f);
    if (!ok) {
      std::printf("Reading error of image tile files\n");
      return 1;
    }
    const int tx = tile % tiles, ty = tile / tiles;
    for (int y = 0; y < h; y++)
      std::memcpy(&full[3 * ((size_t)(ty * h + y) * W + (size_t)tx * w)], &img[3 * (size_t)y * w],
                  sizeof(uint16_t) * 3 * (size_t)w);
    std::remove(path.c_str());
  }
  if (out.empty()) out = std::string(s.file_destination) + "_fromTiles.png";
  make_parent(out);
  const int rc = mbr_write_png(out.c_str(), W, H, full.data(), nullptr, 16);
  if (rc != MBR_OK) {
    std::printf("cannot write %s (status %d)\n", out.c_str(), rc);
    return 1;
  }
  std::printf("image saved: %s (%d x %d from %d tiles)\n", out.c_str(), W, H, tiles * tiles);
  return 0;
}

// Render frames scenes[i] (frame numbers indices[i]) with one
// mbr_render_batch launch into device framebuffers, then copy each frame
// back and finish it as a single render's (finish_image).  0 on success, 1
// on a failure, -1 when the batch does not take these scenes (the caller
// renders them frame by frame).
int render_batch(mbr_ctx *ctx, const std::vector<mbr_scene> &scenes, const std::vector<int> &indices, int precision,
                 const std::string &format) {
  const int n = (int)scenes.size();
  std::vector<mbr_framebuffer> fbs((size_t)n);
  std::vector<void *> blocks((size_t)n, nullptr);
  auto release = [&]() {
    for (void *q : blocks)
      if (q) (void)hipFree(q);
  };
  size_t rays = 0;
  for (int i = 0; i < n; i++) {
    const size_t px = (size_t)scenes[i].image_width * scenes[i].image_height;
    rays += px;
    // rgb x 3 and z (float), alpha and opacity (uint16), colour x 3 (uint8)
    const size_t bytes = px * (4 * sizeof(float) + 2 * sizeof(uint16_t) + 3);
    if (hipMalloc(&blocks[i], bytes) != hipSuccess) {
      std::printf("batch: cannot allocate %zu bytes of device memory\n", bytes);
      release();
      return 1;
    }
    char *q = (char *)blocks[i];
    mbr_framebuffer &fb = fbs[i];
    std::memset(&fb, 0, sizeof fb);
    fb.width = scenes[i].image_width;
    fb.height = scenes[i].image_height;
    for (int c = 0; c < 3; c++, q += px * sizeof(float)) fb.rgb[c] = (float *)q;
    fb.z = (float *)q;
    q += px * sizeof(float);
    fb.alpha = (uint16_t *)q;
    q += px * sizeof(uint16_t);
    fb.opacity = (uint16_t *)q;
    q += px * sizeof(uint16_t);
    for (int c = 0; c < 3; c++, q += px) fb.colour[c] = (uint8_t *)q;
  }
  mbr_batch_opts bo;
  std::memset(&bo, 0, sizeof bo);
  bo.precision = precision;
  bo.replays = 1;
  mbr_stats st;
  mbr_batch_result res;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = mbr_render_batch(ctx, scenes.data(), n, nullptr, fbs.data(), &bo, &st, &res);
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc == MBR_ERR_UNSUPPORTED) {
    std::printf("batch: %s; rendering frame by frame\n", mbr_last_error(ctx));
    release();
    return -1;
  }
  if (rc != MBR_OK) {
    std::printf("batch render failed: %s\n", mbr_last_error(ctx));
    release();
    return 1;
  }
  std::printf("batch of %d frames (%d chains, graph of %d nodes, launch %.3f ms)\n", n, res.slots, res.graph_nodes,
              res.ms_per_launch);
  report(st, sec, rays);
  int status = 0;
  for (int i = 0; i < n && status == 0; i++) {
    Planes pl;
    pl.resize(fbs[i].width, fbs[i].height);
    const size_t px = (size_t)pl.w * pl.h;
    bool ok = true;
    auto get = [&](void *dst, const void *src, size_t bytes) {
      ok = ok && hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
    };
    get(pl.r.data(), fbs[i].rgb[0], px * sizeof(float));
    get(pl.g.data(), fbs[i].rgb[1], px * sizeof(float));
    get(pl.b.data(), fbs[i].rgb[2], px * sizeof(float));
    get(pl.z.data(), fbs[i].z, px * sizeof(float));
    get(pl.cr.data(), fbs[i].colour[0], px);
    get(pl.cg.data(), fbs[i].colour[1], px);
    get(pl.cb.data(), fbs[i].colour[2], px);
    get(pl.alpha.data(), fbs[i].alpha, px * sizeof(uint16_t));
    get(pl.opacity.data(), fbs[i].opacity, px * sizeof(uint16_t));
    if (!ok) {
      std::printf("batch: cannot copy frame %d from the device\n", indices[i]);
      status = 1;
      break;
    }
    status = finish_image(ctx, scenes[i], pl, format, "", indices[i], nullptr);
  }
  release();
  return status;
}

// Frames of an animation, rendered in batches of up to `batch` frames still
// to do (frames whose output exists are skipped first, as render_frame does);
// tiled frames and batch = 1 render frame by frame.
struct AnimationQueue {
  mbr_ctx *ctx;
  int precision, batch;
  std::string format;
  std::vector<mbr_scene> scenes;
  std::vector<int> indices;

  int flush() {
    if (scenes.empty()) return 0;
    int status = scenes.size() > 1 ? render_batch(ctx, scenes, indices, precision, format) : -1;
    if (status < 0) {
      status = 0;
      for (size_t i = 0; i < scenes.size() && status == 0; i++)
        status = render_frame(ctx, scenes[i], precision, format, "", indices[i]);
    }
    scenes.clear();
    indices.clear();
    return status;
  }
  int add(const mbr_scene &fs, int index) {
    if (batch <= 1 || fs.tiles > 1) {
      const int status = flush();
      return status ? status : render_frame(ctx, fs, precision, format, "", index);
    }
    const std::string path = index_filename(fs.file_destination, "png", index);
    if (file_exists(path)) {
      std::printf("Output file (%s) exists. Skip to next file index\n", path.c_str());
      return 0;
    }
    scenes.push_back(fs);
    indices.push_back(index);
    return (int)scenes.size() >= batch ? flush() : 0;
  }
};

}  // namespace

int main(int argc, char **argv) {
  std::vector<std::string> overrides;
  std::string settings, format = "jpg", out;
  int precision = MBR_PREC_FP32_FIXUP;
  int start_frame = -1, end_frame = -1, batch = 16;
  bool flight = false, keyframe = false;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto need = [&](const char *what) -> const char * {
      if (i + 1 >= argc) {
        std::printf("commandline: %s needs an argument\n", what);
        usage();
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "-nogui") continue;
    if (a == "-help" || a == "--help") {
      usage();
      return 0;
    }
    if (a == "-flight") { flight = true; std::printf("commandline: flight mode\n"); continue; }
    if (a == "-keyframe") { keyframe = true; std::printf("commandline: keyframe mode\n"); continue; }
    if (a == "-ip" || a == "-port") {
      std::printf("commandline: %s (network rendering) is not part of this render core\n", a.c_str());
      return 2;
    }
    if (a == "-o") { overrides.push_back(need("-o")); continue; }
    if (a == "-res") {
      int w = 0, h = 0;
      if (std::sscanf(need("-res"), "%dx%d", &w, &h) != 2 || w <= 0 || h <= 0) {
        std::printf("commandline: Argument must be -res WWWWxHHHH\n");
        return 2;
      }
      overrides.push_back("image_width=" + std::to_string(w));
      overrides.push_back("image_height=" + std::to_string(h));
      continue;
    }
    if (a == "-fpk") {
      const int fpk = std::atoi(need("-fpk"));
      if (fpk <= 0) {
        std::printf("commandline: argument must be -fpk NNN\n");
        return 2;
      }
      overrides.push_back("frames_per_keyframe=" + std::to_string(fpk));
      continue;
    }
    if (a == "-start") { start_frame = std::atoi(need("-start")); continue; }
    if (a == "-end") { end_frame = std::atoi(need("-end")); continue; }
    if (a == "-format") {
      format = need("-format");
      if (format != "jpg" && format != "png" && format != "png16" && format != "png16alpha") {
        std::printf("commandline: wrong image format\n");
        return 2;
      }
      continue;
    }
    if (a == "-precision") {
      const std::string p = need("-precision");
      if (p == "fp64") precision = MBR_PREC_FP64;
      else if (p == "fp32") precision = MBR_PREC_FP32_FIXUP;
      else if (p == "fp32march") precision = MBR_PREC_FP32_MARCH;
      else {
        std::printf("commandline: precision must be fp32, fp64 or fp32march\n");
        return 2;
      }
      continue;
    }
    if (a == "-out") { out = need("-out"); continue; }
    if (a == "-batch") {
      batch = std::atoi(need("-batch"));
      if (batch < 1) {
        std::printf("commandline: argument must be -batch N (N >= 1)\n");
        return 2;
      }
      continue;
    }
    if (!a.empty() && a[0] == '-') {
      std::printf("commandline: wrong parameters\n");
      usage();
      return 2;
    }
    settings = a;
  }
  if (settings.empty() || (flight && keyframe)) {
    usage();
    return 2;
  }
  std::vector<const char *> ov;
  for (const std::string &o : overrides) ov.push_back(o.c_str());
  mbr_scene s;
  int rc = mbr_load_settings(settings.c_str(), ov.empty() ? nullptr : ov.data(), (int)ov.size(), &s);
  if (rc != MBR_OK) {
    std::printf("cannot load settings %s (status %d)\n", settings.c_str(), rc);
    return 1;
  }
  mbr_ctx *ctx = nullptr;
  rc = mbr_init(nullptr, &ctx);
  if (rc != MBR_OK) {
    std::printf("mbr_init failed (status %d): no usable GPU\n", rc);
    return 1;
  }
  // the random aux lights are placed once, from the loaded settings, and
  // keep their places over an animation; each frame rewrites the predefined
  // ones (src/Render3D.cpp:1990-1993, :2305)
  if (s.aux_light_number > 0) {
    std::vector<mbr_light> tab((size_t)s.aux_light_number + 4);
    int count = 0;
    rc = mbr_place_lights(&s, tab.data(), (int)tab.size(), &count);
    if (rc == MBR_OK) rc = mbr_set_lights(ctx, tab.data(), count);
    if (rc != MBR_OK) {
      std::printf("aux light placement failed (status %d)\n", rc);
      mbr_destroy(ctx);
      return 1;
    }
    std::printf("%d aux lights placed\n", s.aux_light_number);
  }
  int status = 0;
  if (!flight && !keyframe) {
    // a still image: frame number -start (0 by default)
    const int index = start_frame > 0 ? start_frame : 0;
    s.frame_no = index;
    status = render_frame(ctx, s, precision, format, out, index);
    mbr_destroy(ctx);
    return status;
  }
  // animation frame range (src/Render3D.cpp:2088-2099)
  int first = s.start_frame, last = s.end_frame;
  if (start_frame > 0) first = start_frame;
  if (end_frame >= 0 && end_frame < 99999) last = end_frame;
  if (keyframe) {
    // keyframes: file_keyframes + %05d + .fract while they exist (:2020-2068)
    std::vector<mbr_scene> keys;
    for (int k = 0;; k++) {
      const std::string path = index_filename(s.file_keyframes, "fract", k);
      if (!file_exists(path)) break;
      mbr_scene ks;
      rc = mbr_load_settings(path.c_str(), nullptr, 0, &ks);
      if (rc != MBR_OK) {
        std::printf("cannot load keyframe %s (status %d)\n", path.c_str(), rc);
        mbr_destroy(ctx);
        return 1;
      }
      keys.push_back(ks);
    }
    std::printf("Found %d keyframes\n", (int)keys.size());
    if (keys.empty()) {
      mbr_destroy(ctx);
      return 1;
    }
    const int fpk = s.frames_per_keyframe > 0 ? s.frames_per_keyframe : 1;
    if (last > (int)keys.size() * fpk) last = (int)keys.size() * fpk;
    AnimationQueue queue{ctx, precision, batch, format, {}, {}};
    for (int index = first; index < last && status == 0; index++) {
      mbr_scene fs = s;
      rc = mbr_keyframe_scene(keys.data(), (int)keys.size(), fpk, index, &fs);
      if (rc != MBR_OK) {
        status = 1;
        break;
      }
      fs.frame_no = index;
      std::printf("Frame: %d, Keyframe %f\n", index, (double)index / fpk);
      status = queue.add(fs, index);
    }
    if (status == 0) status = queue.flush();
  } else {
    // flight path: one camera per line (:2262-2274), skipping to the first frame
    FILE *path = std::fopen(s.file_animation_path, "r");
    if (!path) {
      std::printf("cannot open flight path %s\n", s.file_animation_path);
      mbr_destroy(ctx);
      return 1;
    }
    mbr_scene fs = s;
    auto read_camera = [&](mbr_scene &t) {
      return std::fscanf(path, "%lf %lf %lf %lf %lf %lf", &t.vp[0], &t.vp[1], &t.vp[2], &t.alpha, &t.beta,
                         &t.gamma) == 6;
    };
    bool ok = true;
    for (int i = 0; i < first && ok; i++) ok = read_camera(fs);
    AnimationQueue queue{ctx, precision, batch, format, {}, {}};
    for (int index = first; index < last && ok && status == 0; index++) {
      if (!read_camera(fs)) break;  // the path ended: the animation ends
      fs.frame_no = index;
      std::printf("Frame: %d\n", index);
      status = queue.add(fs, index);
    }
    if (status == 0) status = queue.flush();
    std::fclose(path);
  }
  mbr_destroy(ctx);
  return status;
}
