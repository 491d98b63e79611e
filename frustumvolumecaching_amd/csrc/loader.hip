// loader.hip -- the model file layout of the reference's nerfacc export
// (load_nerfacc_export, src/loader.cu:232-285; parse_helper.h:25-63):
//
//   <dir>/config.json         "scene" {grid_resolution, grid_nlvl, aabb[6],
//                             optional contraction.aabb[6], is_open_gl,
//                             alpha_thre, cone_angle, stepsize, near, far},
//                             "mlp_base" {encoding, network, n_output},
//                             "mlp_head" {encoding, network}
//   <dir>/pos_enc.dat         fp16 hash-grid table (optional: when absent the
//                             table follows the weights in mlp_base.dat)
//   <dir>/mlp_base.dat        fp16 density / latent MLP weights
//   <dir>/mlp_head.dat        fp16 colour MLP weights
//   <dir>/mlp_first_head.dat  fp16 first-head weights (split models: config
//                             key "mlp_first_head", loader.cu:251-262)
//   <dir>/occupancy_grid.dat  uint8 [grid_nlvl][res][res][res], > 0 occupied
//
// Every .dat file is raw little-endian data without a header (buffer.h:105-
// 121).  Host code only; the model is then created by fvc_model_create.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "fvc_internal.h"
#include "json_min.h"

extern "C" int fvc_stepsize_setup(fvc_stepsize_info* s, const fvc_scene_info* sc);

namespace fvc {

static std::string join(const std::string& dir, const char* name) {
  if (dir.empty() || dir.back() == '/') return dir + name;
  return dir + "/" + name;
}

// size in bytes, -1 when the file cannot be opened
static int64_t file_bytes(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) return -1;
  return (int64_t)f.tellg();
}

static void read_exact(const std::string& path, void* dst, int64_t bytes) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(FVC_E_INVALID, "cannot open " + path);
  f.read(static_cast<char*>(dst), bytes);
  if (f.gcount() != bytes) throw Error(FVC_E_INVALID, "short read from " + path);
}

static float f3(const json::Value& a, size_t i) { return (float)a.at(i).number(); }

// SH degree of the head's direction encoding: a SphericalHarmonics encoding,
// directly or as the first nested one of a Composite (nerf_network.h:104-108)
static int sh_degree_of(const json::Value& enc) {
  if (enc.has("nested")) return (int)enc.at("nested").at(0).number_or("degree", 4);
  return (int)enc.number_or("degree", 4);
}

void nerfacc_read_config(const std::string& dir, NerfaccExport& ex) {
  std::ifstream cf(join(dir, "config.json"));
  if (!cf) throw Error(FVC_E_INVALID, "Could not load config file " + join(dir, "config.json"));
  std::stringstream ss;
  ss << cf.rdbuf();
  json::Value cfg;
  try {
    cfg = json::parse(ss.str());
  } catch (const std::exception& e) {
    throw Error(FVC_E_INVALID, e.what());
  }
  try {
    // parse(SceneInfo&) (parse_helper.h:25-54)
    const json::Value& sc = cfg.at("scene");
    fvc_scene_info& s = ex.scene;
    s = fvc_scene_info{};
    s.grid_resolution = (int)sc.at("grid_resolution").number();
    s.grid_nlvl = (int)sc.at("grid_nlvl").number();
    s.vals_per_lvl = s.grid_resolution * s.grid_resolution * s.grid_resolution;
    s.normalized = 0;
    const json::Value& aabb = sc.at("aabb");
    s.aabb_from = fvc_float3{f3(aabb, 0), f3(aabb, 1), f3(aabb, 2)};
    s.aabb_to = fvc_float3{f3(aabb, 3), f3(aabb, 4), f3(aabb, 5)};
    if (sc.has("contraction") && sc.at("contraction").has("aabb")) {
      const json::Value& c = sc.at("contraction").at("aabb");
      s.contraction_type = FVC_CONTRACTION_WARP_AABB_LINF;
      s.contraction_aabb_from = fvc_float3{f3(c, 0), f3(c, 1), f3(c, 2)};
      s.contraction_aabb_to = fvc_float3{f3(c, 3), f3(c, 4), f3(c, 5)};
    } else {
      s.contraction_type = FVC_CONTRACTION_AABB;
    }
    s.is_open_gl = sc.at("is_open_gl").boolean() ? 1 : 0;
    s.alpha_thre = (float)sc.at("alpha_thre").number();
    // parse(StepsizeInfo&) (parse_helper.h:57-63) + setAdditionalInfo
    fvc_stepsize_info& st = ex.step;
    st = fvc_stepsize_info{};
    st.cone_angle = (float)sc.at("cone_angle").number();
    st.stepsize = (float)sc.at("stepsize").number();
    st.near_ = (float)sc.at("near").number();
    st.far_ = (float)sc.at("far").number();
    if (fvc_stepsize_setup(&st, &s) != FVC_OK) throw Error(FVC_E_INVALID, "stepsize setup failed");
    // NerfNetwork::init (nerf_network.h:88-118): tcnn HashGrid + FullyFusedMLPs
    const json::Value& base = cfg.at("mlp_base");
    const json::Value& enc = base.at("encoding");
    const json::Value& head = cfg.at("mlp_head");
    fvc_net_desc& n = ex.net;
    n = fvc_net_desc{};
    n.n_levels = (int)enc.at("n_levels").number();
    n.n_features_per_level = (int)enc.number_or("n_features_per_level", 2);
    n.log2_hashmap_size = (int)enc.at("log2_hashmap_size").number();
    n.base_resolution = (int)enc.at("base_resolution").number();
    n.per_level_scale = (float)enc.at("per_level_scale").number();
    n.base_width = (int)base.at("network").at("n_neurons").number();
    n.base_n_hidden_layers = (int)base.at("network").at("n_hidden_layers").number();
    n.base_n_output = (int)base.at("n_output").number();
    n.head_width = (int)head.at("network").at("n_neurons").number();
    n.head_n_hidden_layers = (int)head.at("network").at("n_hidden_layers").number();
    if (head.has("latent_offset") && (int)head.at("latent_offset").number() == 1) {
      // the synthetic separate-density layout ("latent_offset": 1 is this
      // library's key, not the reference's): log sigma in row 0, the 8
      // latents after it, the density kept apart in the cache
      n.latent_offset = 1;
      n.n_latents = 8;
      n.first_latent_is_density = 0;
    } else {
      // NerfNetwork::init (nerf_network.h:88-118): m_n_latents = n_output
      // and m_first_latent_is_density stays true, so latent_offset() = 0
      // (nerf_network.h:93, 334-342): the head reads rows 0..n_output-1,
      // log sigma first; eval picks DensityIntermediates (eval.cu:216)
      n.latent_offset = 0;
      n.n_latents = n.base_n_output;
      n.first_latent_is_density = 1;
    }
    // head input: the latents, SH, then padding with 1.0 up to the head's
    // 32 inputs (nerf_network.h:108-109, 248-253); without "n_input" the
    // reference leaves those rows as the buffer holds them, here they are
    // the padding
    if (head.has("n_input") && !cfg.has("mlp_first_head") &&
        (int)head.at("n_input").number() != n.n_latents + 16)
      throw Error(FVC_E_UNSUPPORTED, "mlp_head n_input must be the latents + 16 SH inputs");
    n.sh_degree = sh_degree_of(head.at("encoding"));
    n.precision = FVC_PREC_FP16;
    if (cfg.has("mlp_first_head")) {
      // NerfNetwork::init_viewdep (nerf_network.h:120-158)
      const json::Value& fh = cfg.at("mlp_first_head");
      n.split_head = 1;
      n.first_head_width = (int)fh.at("network").at("n_neurons").number();
      n.first_head_n_hidden_layers = (int)fh.at("network").at("n_hidden_layers").number();
      n.first_head_n_input = (int)fh.at("n_input").number();
      n.first_latent_is_density = fh.has("first_latent_is_density") ? (fh.at("first_latent_is_density").boolean() ? 1 : 0) : 1;
      n.first_sh_degree = sh_degree_of(fh.at("encoding"));
      n.n_latents = (int)fh.at("n_output").number();
      n.latent_offset = n.first_latent_is_density ? 0 : 1;
      n.freq_encoding_degree = (int)head.number_or("freq_encoding_degree", 0);
      // the custom head's input width (my_fully_fused_mlp.cu:1300-1308)
      if (head.has("n_input") && (int)head.at("n_input").number() != vd_head_inputs(n))
        throw Error(FVC_E_UNSUPPORTED, "mlp_head n_input does not match the split head's inputs");
    }
  } catch (const Error&) {
    throw;
  } catch (const std::exception& e) {
    throw Error(FVC_E_INVALID, e.what());
  }
  GridParams g;
  int64_t n_pos = 0, n_base = 0, n_head = 0, n_first = 0;
  if (!net_layout(ex.net, g, n_pos, n_base, n_head, &n_first))
    throw Error(FVC_E_UNSUPPORTED, "network description in config.json is not supported");
  ex.n_pos = n_pos;
  ex.n_base = n_base;
  ex.n_head = n_head;
  ex.n_first = n_first;
  ex.n_occ = (int64_t)ex.scene.vals_per_lvl * ex.scene.grid_nlvl;
  const int64_t b_pos = file_bytes(join(dir, "pos_enc.dat"));
  const int64_t b_base = file_bytes(join(dir, "mlp_base.dat"));
  const int64_t b_head = file_bytes(join(dir, "mlp_head.dat"));
  const int64_t b_occ = file_bytes(join(dir, "occupancy_grid.dat"));
  ex.pos_separate = b_pos >= 0;
  if (b_base < 0 || b_head < 0 || b_occ < 0) throw Error(FVC_E_INVALID, "missing .dat file in " + dir);
  const int64_t want_base = 2 * (ex.pos_separate ? n_base : n_base + n_pos);
  if ((ex.pos_separate && b_pos != 2 * n_pos) || b_base != want_base || b_head != 2 * n_head || b_occ != ex.n_occ)
    throw Error(FVC_E_INVALID, "parameter file sizes do not match config.json");
  if (ex.net.split_head && file_bytes(join(dir, "mlp_first_head.dat")) != 2 * n_first)
    throw Error(FVC_E_INVALID, "mlp_first_head.dat missing or its size does not match config.json");
}

void nerfacc_read_data(const std::string& dir, const NerfaccExport& ex, uint16_t* pos, uint16_t* base, uint16_t* head,
                       uint8_t* occ, uint16_t* first) {
  if (ex.net.split_head) {
    if (!first) throw Error(FVC_E_INVALID, "split model: mlp_first_head.dat needs a buffer");
    read_exact(join(dir, "mlp_first_head.dat"), first, 2 * ex.n_first);
  }
  if (ex.pos_separate) {
    read_exact(join(dir, "pos_enc.dat"), pos, 2 * ex.n_pos);
    read_exact(join(dir, "mlp_base.dat"), base, 2 * ex.n_base);
  } else {
    // set_params(params_mlp_base, ...) (nerf_network.h:279-282): the table
    // follows the MLP weights in mlp_base.dat
    std::vector<uint16_t> both((size_t)(ex.n_base + ex.n_pos));
    read_exact(join(dir, "mlp_base.dat"), both.data(), 2 * (ex.n_base + ex.n_pos));
    std::memcpy(base, both.data(), 2 * ex.n_base);
    std::memcpy(pos, both.data() + ex.n_base, 2 * ex.n_pos);
  }
  read_exact(join(dir, "mlp_head.dat"), head, 2 * ex.n_head);
  read_exact(join(dir, "occupancy_grid.dat"), occ, ex.n_occ);
}

}  // namespace fvc
