/* gp5_nn.h -- kernels of the learned warm start (a 3D convolutional encoder-
 * decoder that predicts droplet ground states).  No counterpart in the
 * reference, which has no learned component: these serve the callers that
 * seed imaginary-time ground states (the batch API of gp5.h) and the training
 * loop that fits the network to ground states computed by that solver.
 *
 * Tensors are plain device pointers.  Activations are channels-last
 * [batch][z][y][x][channel] cubes of side D, in bfloat16 (uint16 bit patterns)
 * for the matrix-core inputs and in float for gradients and outputs.  Weights
 * of a 3x3x3 "same" convolution are float masters W[tap][ci][co]
 * (tap = 9 dz + 3 dy + dx over offsets -1..1) with bfloat16 copies in two
 * layouts for the forward and data-gradient kernels.  Channel counts are
 * multiples of 16.  Kernels and copies go to the calling thread's stream
 * (gp5_nn_set_stream; the null stream until set); every call returns GP5_OK or
 * an error code with gp5_last_error() set (include/gp5.h).  The library's own
 * scratch (split convolutions, the loss, the bias-gradient partials of
 * gp5_nn_bias_grad and gp5_nn_act_bwd_cast_bias) is one per process: calls that
 * use it must not run on two streams at once.  The weight gradients use only
 * the caller's workspace. */
#ifndef GP5_NN_H
#define GP5_NN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* device memory and transfers (upload and download return when done; copy and
 * zero are queued on the thread's stream) */
int gp5_nn_malloc(void **p, size_t bytes);
int gp5_nn_free(void *p);
int gp5_nn_upload(void *dst, const void *host, size_t bytes);
int gp5_nn_download(void *host, const void *src, size_t bytes);
int gp5_nn_copy(void *dst, const void *src, size_t bytes); /* device to device */
int gp5_nn_zero(void *p, size_t bytes); /* fill kernel, not a memset node */
int gp5_nn_sync(void); /* waits for every stream of the device */

/* Streams and events, so that independent passes of a training step (a layer's
 * weight gradient beside the data gradients that follow it) run side by side.
 * Streams are non-blocking (they do not wait for the null stream) at the default
 * priority; events carry no timing.  gp5_nn_set_stream(NULL) returns the calling
 * thread to the null stream.  gp5_nn_event_record marks the work queued so far
 * on a stream; gp5_nn_stream_wait makes later work on a stream wait for it. */
int gp5_nn_stream_create(void **stream);
int gp5_nn_stream_destroy(void *stream);
int gp5_nn_set_stream(void *stream);
int gp5_nn_event_create(void **event);
int gp5_nn_event_destroy(void *event);
int gp5_nn_event_record(void *event, void *stream);
int gp5_nn_stream_wait(void *stream, void *event);

/* Y = act(conv(X, W) + bias) (+ skip).  Implicit GEMM on bf16 MFMA tiles with
 * float accumulation.  Wk: bf16 [27][Cout][Cin] (forward layout); with
 * mirror = 1 the taps are read mirrored, which with the [27][ci][co] layout
 * passed as Wk computes the data gradient (Cin/Cout then name the gradient's
 * input/output channels).  act: 0 identity, 1 leaky ReLU (slope 0.1).
 * Outputs (each may be NULL): A bf16 = act(...); Y bf16 = A + skip;
 * Yf float = act(...) (for the last layer and for gradients). */
int gp5_nn_conv3d(const void *X, const void *Wk, const float *bias, void *A, const void *skip, void *Y, float *Yf,
                  int B, int D, int Cin, int Cout, int mirror, int act);
/* The float output of the convolution (no bias, no activation; mirror as in
 * gp5_nn_conv3d) summed over the 2x2x2 children of each cell of side D/2: Yh
 * (B*(D/2)^3*Cout floats) equals gp5_nn_conv3d into Yf followed by
 * gp5_nn_pool2_f32(Yf, Yh, B, D, Cout, 8), bit for bit, without writing Yf (the
 * data gradient of a layer that reads the nearest upsampling of its input).
 * Cin = 16 with Cout a multiple of 32, or Cin = 32; D = 32, 64 or 128; other
 * shapes return GP5_ERR_ARG. */
int gp5_nn_conv3d_pool(const void *X, const void *Wk, float *Yh, int B, int D, int Cin, int Cout, int mirror);
/* The data gradient of a layer (gp5_nn_conv3d with mirrored taps, no bias or
 * activation: dY of Cin channels, Wk its [27][ci][co] copy, Yf float of Cout
 * channels) followed by the backward of the leaky ReLU of the layer below, whose
 * output is Ab (bf16, Cout channels): H = bf16(Yf act'(Ab)) and the block partials
 * of that layer's bias gradient in part (gp5_nn_bias_part_floats(Cout) floats,
 * finished by gp5_nn_bias_final_multi), as gp5_nn_act_bwd_part(Ab, Yf, NULL, 0,
 * H, part, ...) would give after the convolution (Yf and H bit for bit; the
 * partials are the z-sliding kernel's per-workgroup sums where it takes the shape
 * with one 16-channel output block per wave, one pass; otherwise the two passes
 * run).  Cout / 4 must divide 256. */
int gp5_nn_conv3d_act_bwd(const void *dY, const void *Wk, const void *Ab, float *Yf, void *H, float *part, int B,
                          int D, int Cin, int Cout);
/* gp5_nn_conv3d_act_bwd for the output layer's gradient, given as channel 0 alone
 * (dY0: one bf16 per position, gp5_nn_mse_out0c's H0; the loss reads channel 0, so
 * the other Cin - 1 channels are zero): the data gradient v[p][n] = sum_tap
 * Wk[26 - tap][n][0] dY0[p + off(tap)] goes out as Gd = bf16(v) (Cout channels),
 * H = bf16(v act'(Ab)) and the bias partials of v act'(Ab) as gp5_nn_conv3d_act_bwd
 * gives them.  One 16x16x32 MFMA per 16 positions and 16 channels with the 27 taps
 * as K; v agrees with gp5_nn_conv3d_act_bwd's Yf to float rounding.  D a power of
 * two in [16, 128], Cout a multiple of 16. */
int gp5_nn_conv3d_act_bwd_out0(const void *dY0, const void *Wk, const void *Ab, void *Gd, void *H, float *part, int B,
                               int D, int Cin, int Cout);
/* The output layer's form: no activation, and the float output Yf0 keeps only
 * channel 0 (one float per position, B*D^3 floats); the other channels are
 * computed (the MFMA tiles are 16 wide) and dropped. */
int gp5_nn_conv3d_out0(const void *X, const void *Wk, const float *bias, float *Yf0, int B, int D, int Cin, int Cout);
/* The same convolution of the nearest upsampling of Xh (side D/2, Cin channels)
 * without forming it: output side D, a power of two in [8, 128]. */
int gp5_nn_conv3d_up(const void *Xh, const void *Wk, const float *bias, void *A, const void *skip, void *Y, float *Yf,
                     int B, int D, int Cin, int Cout, int mirror, int act);
/* dW[tap][ci][co] = sum over positions of X[p + off(tap)][ci] dY[p][co]
 * (bf16 inputs, float result, summed in a fixed order); ws: device workspace of
 * at least gp5_nn_wgrad_ws(B, D, Cin, Cout) bytes (per-workgroup partial sums;
 * for cube sides outside 16..128 also channel-major copies of X and dY). */
size_t gp5_nn_wgrad_ws(int B, int D, int Cin, int Cout);
int gp5_nn_conv3d_wgrad(const void *X, const void *dY, float *dW, float *ws, int B, int D, int Cin, int Cout);
/* The weight gradient of gp5_nn_conv3d_up: X is Xh (side D/2), read upsampled;
 * D a multiple of 8; the same workspace size. */
int gp5_nn_conv3d_wgrad_up(const void *Xh, const void *dY, float *dW, float *ws, int B, int D, int Cin, int Cout);
/* gp5_nn_conv3d_wgrad for a dY that is nonzero in channel 0 only (the output layer's
 * gradient rows from gp5_nn_mse_out0): dW[t][ci][0] from channel 0, the other output
 * channels' gradients 0.  With 16 input channels and D = 32, 64 or 128 the 27 taps
 * are the MFMA's N dimension (one 16x16x32 MFMA per 16 taps and 32 positions instead
 * of 27); otherwise the full weight gradient runs (its other columns come out 0). */
int gp5_nn_conv3d_wgrad_out0(const void *X, const void *dY, float *dW, float *ws, int B, int D, int Cin, int Cout);
/* gp5_nn_conv3d_wgrad_out0 with dY given as channel 0 alone (dY0: one bf16 per
 * position); the same result bit for bit.  Cin = 16, Cout a multiple of 16 and
 * D = 32, 64 or 128 (other shapes are refused). */
int gp5_nn_conv3d_wgrad_out0c(const void *X, const void *dY0, float *dW, float *ws, int B, int D, int Cin, int Cout);
/* db[c] = sum over positions of dY[p][c] (float). */
int gp5_nn_bias_grad(const float *dY, float *db, long long P, int C);

/* 2x2x2 average pool (D -> D/2) and nearest upsample (D -> 2D), times scale */
int gp5_nn_pool2_bf16(const void *X, void *Y, int B, int D, int C, float scale);
int gp5_nn_up2_bf16(const void *X, void *Y, int B, int D, int C, float scale);
int gp5_nn_pool2_f32(const float *X, float *Y, int B, int D, int C, float scale);
int gp5_nn_up2_f32(const float *X, float *Y, int B, int D, int C, float scale);
/* dPre = dA * (A > 0 ? 1 : 0.1): backward of the leaky ReLU from its output */
int gp5_nn_act_bwd(const void *A, const float *dA, float *dPre, long long n);
int gp5_nn_add_f32(float *dst, const float *src, long long n);
/* dPre = dA act'(A) (float) and H = bf16(dPre) in one pass; n a multiple of 4;
 * dPre may be dA. */
int gp5_nn_act_bwd_cast(const void *A, const float *dA, float *dPre, void *H, long long n);
/* H = bf16(dA act'(A)) for P positions of C channels (C / 4 divides 256) and the
 * bias gradient db[c] = sum over positions of dA act'(A) (float), in one pass. */
int gp5_nn_act_bwd_cast_bias(const void *A, const float *dA, void *H, float *db, long long P, int C);
/* The same with the incoming gradient dA + scale * G at the parent cell, G of
 * side D/2 (A, dA, H of side D, B samples): gp5_nn_up2_add_f32 into dA followed
 * by gp5_nn_act_bwd_cast_bias, bit for bit, without writing dA. */
int gp5_nn_up2_add_act_bwd_cast_bias(const void *A, const float *dA, const float *G, float scale, void *H, float *db,
                                     int B, int D, int C);
/* The activation backward of the two calls above (G may be NULL) with the bias
 * gradient left as block partials in `part` (device, gp5_nn_bias_part_floats(C)
 * floats), so that several layers' bias gradients finish in one launch of
 * gp5_nn_bias_final_multi (same values as the calls above, bit for bit). */
size_t gp5_nn_bias_part_floats(int C);
int gp5_nn_act_bwd_part(const void *A, const float *dA, const float *G, float scale, void *H, float *part, int B, int D,
                        int C);
/* gp5_nn_act_bwd_part with G (required) and dA in bf16 (gp5_nn_conv3d_act_bwd_out0's
 * Gd): the same arithmetic on the bf16 values. */
int gp5_nn_act_bwd_part_bf16(const void *A, const void *dA, const float *G, float scale, void *H, float *part, int B,
                             int D, int C);
typedef struct gp5_nn_bias_seg {
    const float *part; /* partials of gp5_nn_act_bwd_part */
    float *db;         /* C floats: the bias gradient */
    int C;
} gp5_nn_bias_seg;
#define GP5_NN_BIAS_MAX_SEGS 16
int gp5_nn_bias_final_multi(const gp5_nn_bias_seg *segs, int nseg);
/* Y (side 2D) += scale * X at the parent cell (the adjoint of pooling added into
 * Y); C a multiple of 4. */
int gp5_nn_up2_add_f32(const float *X, float *Y, int B, int D, int C, float scale);
int gp5_nn_cast_bf16(const float *src, void *dst, long long n);
/* Mean squared error of y = base + Yf[p][0] against target over P positions:
 * dYf[p][c] = 2 (y - t) / P for c = 0 and 0 otherwise; *loss = mean. */
int gp5_nn_mse(const float *Yf, int C, const float *base, const float *target, float *dYf, long long P,
               double *loss);
/* The same, also writing H = bf16(dYf) (C a multiple of 8); dYf may be NULL (not
 * written); db, if not NULL, receives the output layer's bias gradient, the sum
 * over positions of the gradient rows (channel 0; the others are 0). */
int gp5_nn_mse_bf16(const float *Yf, int C, const float *base, const float *target, float *dYf, void *H, float *db,
                    long long P, double *loss);
/* gp5_nn_mse_bf16 for a prediction y = base + y0 with y0 one float per position
 * (gp5_nn_conv3d_out0's output); H rows have C channels (a multiple of 8),
 * channel 0 the gradient 2 (y - t) / P, the others 0; db (C floats) as above. */
int gp5_nn_mse_out0(const float *y0, const float *base, const float *target, void *H, int C, float *db, long long P,
                    double *loss);
/* gp5_nn_mse_out0 that adds the loss to loss_acc[0] (a device double) instead of
 * copying it to the host: no synchronisation, so a training loop reads the sum of
 * its losses when it needs them (one loss per call, added in call order). */
int gp5_nn_mse_out0_acc(const float *y0, const float *base, const float *target, void *H, int C, float *db,
                        long long P, double *loss_acc);
/* gp5_nn_mse_out0 that writes channel 0 of the gradient alone (H0: one bf16 per
 * position, the values gp5_nn_mse_out0 writes to channel 0 of its rows); db
 * receives C floats as there.  The loss goes to *loss (host) and/or is added to
 * loss_acc[0] (device), whichever is not NULL. */
int gp5_nn_mse_out0c(const float *y0, const float *base, const float *target, void *H0, int C, float *db, long long P,
                     double *loss, double *loss_acc);
/* Droplet input channels (one cube of side D on the scaled box [-ub, ub]^3, C-ordered
 * [D][D][D][16] bf16): 0 = 0.5 (1 - tanh((r - 1) s)), 1 = a0 u_i^2 + a1 u_j^2 + a2 u_k^2,
 * 2..6 = c[0..4], 7 = r / ub, 8..15 = 0; base (float, D^3) = channel 0.  float32
 * arithmetic in the host construction's order, rounded to nearest even. */
int gp5_nn_droplet_features(void *X, float *base, int D, double ub, float s, const float a[3], const float c[5]);
/* y[p] = base[p] + Yf[p * C]: the prediction (baseline plus output channel 0). */
int gp5_nn_predict_out(const float *Yf, int C, const float *base, float *y, long long P);
/* Adam step on n float parameters (t: 1-based step count). */
int gp5_nn_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2, float eps,
                int t);
/* bf16 copies of W[27][Cin][Cout]: Wf [27][Cout][Cin] (forward), Wb = same layout as W. */
int gp5_nn_relayout(const float *W, void *Wf, void *Wb, int Cin, int Cout);
/* One tensor of gp5_nn_adam_multi: p, g, m, v as in gp5_nn_adam; when wf or wb is
 * non-null the tensor is W[27][cin][cout] and its updated values also go to the
 * bf16 copies of gp5_nn_relayout (either may be null). */
typedef struct gp5_nn_adam_seg {
    float *p;
    const float *g;
    float *m, *v;
    void *wf, *wb;
    long long n;
    int cin, cout;
} gp5_nn_adam_seg;
#define GP5_NN_ADAM_MAX_SEGS 16
/* Adam step on up to GP5_NN_ADAM_MAX_SEGS tensors in one launch, with the bf16
 * copies of the weight tensors refreshed in the same pass: element by element
 * the arithmetic of gp5_nn_adam followed by gp5_nn_relayout, so the results
 * are identical to those calls (which take two launches per tensor). */
int gp5_nn_adam_multi(const gp5_nn_adam_seg *segs, int nseg, float lr, float b1, float b2, float eps, int t);

#ifdef __cplusplus
}
#endif

#endif /* GP5_NN_H */
