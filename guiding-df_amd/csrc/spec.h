// Plain-old-data description of one network of the guide (encoding + MLP +
// parameter/fragment layout).  Built on the host from the JSON configuration
// and passed by value to every kernel, so host and device share one layout.
#pragma once
#include <stdint.h>

#define DFG_MAX_PARTS 6
#define DFG_MAX_LEVELS 24
#define DFG_MAX_IN 16
#define DFG_MAX_HIDDEN 8
#define DFG_WIDTH 64          // hidden width of every MLP (FullyFusedMLP<64>)
#define DFG_MAX_ENC 128       // largest encoded width after padding to 32
#define DFG_TILE 64           // rows (queries / samples) per wavefront tile

namespace dfg {

enum EncType { ENC_IDENTITY = 0, ENC_ONEBLOB = 1, ENC_SH = 2, ENC_GRID = 3, ENC_TRIWAVE = 4, ENC_FREQUENCY = 5 };
enum Interp { INTERP_NEAREST = 0, INTERP_LINEAR = 1, INTERP_SMOOTHSTEP = 2 };
enum GridType { GRID_DENSE = 0, GRID_HASH = 1, GRID_TILED = 2 };

// One nested encoding of a Composite (encodings/composite.h:136-215).
struct EncPart {
    int type;
    int dim_begin;   // first raw input dimension it reads
    int n_dims;      // raw input dimensions it reads
    int out_begin;   // first encoded column of its block (pads included)
    int n_out;       // encoded columns written
    int n_pad;       // padding columns (value 1.0)
    int pad_front;   // 1: padding precedes the outputs (spherical harmonics), 0: follows
    int param;       // OneBlob n_bins | SH degree | TriangleWave/Frequency n_frequencies | grid F
    float id_scale;  // Identity scale / offset
    float id_offset;
};

// Multiresolution grid (encodings/grid.h:943-1025).  At most one per net.
struct GridSpec {
    int present;
    int n_levels;
    int n_feat;        // features per level F
    int interp;        // Interp
    int type;          // GridType
    int base_res;
    float per_level_scale;
    uint32_t param_begin;             // first grid parameter inside the net's parameter vector
    uint32_t offset[DFG_MAX_LEVELS + 1];  // level offsets in table entries
    uint32_t res[DFG_MAX_LEVELS];     // vertices per axis
    float scale[DFG_MAX_LEVELS];      // grid_scale(level)
    // Per-level closed form of grid_index (precomputed on the host from the
    // reference's uint32 stride loop): dense index x + y*s1 + z*s2 (uint32
    // wrap, s = 0 for dims the loop never reached) or the CoherentPrime hash,
    // then "% hsize" by mask (power of two) or by the Granlund-Montgomery
    // multiply-shift for invariant divisors.
    uint32_t s1[DFG_MAX_LEVELS];
    uint32_t s2[DFG_MAX_LEVELS];
    uint32_t hashed[DFG_MAX_LEVELS];
    uint32_t mod_mask[DFG_MAX_LEVELS];   // hsize - 1 when hsize is a power of two, else 0
    uint32_t mod_mul[DFG_MAX_LEVELS];    // Granlund-Montgomery m'
    uint32_t mod_shift[DFG_MAX_LEVELS];  // l - 1
};

struct NetSpec {
    int in_dims;       // raw input dimensions
    int n_parts;
    EncPart parts[DFG_MAX_PARTS];
    GridSpec grid;
    int enc_width;     // encoded width padded to 16 (the reference MLP input width)
    int k_steps;       // ceil(enc_width / 32): K-steps of the first MFMA layer
    int grid_ksteps;   // leading K-steps made of grid features alone (0: none)
    int n_hidden;      // hidden layers (n_hidden_layers)
    int n_out;         // real outputs (bins, or 3 for RGB)
    int n_out_pad;     // outputs padded to 16
    int n_layers;      // n_hidden + 1 matrices
    uint32_t layer_off[DFG_MAX_HIDDEN + 2];  // matrix offsets inside the parameter vector
    uint32_t n_matrix_params;
    uint32_t n_params;
    // MFMA fragment packs (in halves): forward packs of W_l, backward packs of W_l^T
    uint32_t fwd_off[DFG_MAX_HIDDEN + 2];
    uint32_t bwd_off[DFG_MAX_HIDDEN + 2];
    uint32_t pack_size;  // halves of one complete set (forward + backward)
};

}  // namespace dfg
