// TEST INFRASTRUCTURE: a CPU stand-in for libtfim.so, linked only into the
// test build of the engine (tests/test_engine_host_logic.py) so the search
// logic can be checked where there is no GPU.  The product never loads it.
#include "tfim.h"
#include <algorithm>
#include <cstring>
#include <vector>
#include <string>
#include <cstdlib>
struct tfim_ctx { std::vector<uint64_t> bm; int64_t W=0, n_bits=0; int32_t n_items=0, n_cols=0; std::string err;
                  std::vector<int32_t> p_ids; std::vector<int64_t> p_off; int32_t p_max=-1; };
extern "C" {
int tfim_open(int, tfim_ctx **o) { *o = new tfim_ctx(); return 0; }
void tfim_close(tfim_ctx *c) { delete c; }
int tfim_open_shared(tfim_ctx *p, tfim_ctx **o) { *o = new tfim_ctx(*p); return 0; }
const char *tfim_last_error(const tfim_ctx *c) { return c->err.c_str(); }
int tfim_device_info(tfim_ctx *, char *n, int l, int32_t *cu, int64_t *h, int32_t *clk) { snprintf(n, l, "mock"); if(cu)*cu=0; if(h)*h=0; if(clk)*clk=0; return 0; }
int tfim_config(tfim_ctx *, const char *, int64_t) { return 0; }
int tfim_load_bitmaps(tfim_ctx *c, const uint64_t *rows, int32_t n, int64_t nb, int64_t w, int32_t nc) {
  c->W = (nb+63)/64; c->bm.assign((size_t)n*c->W, 0); for (int i=0;i<n;i++) memcpy(&c->bm[(size_t)i*c->W], rows+(size_t)i*w, c->W*8);
  c->n_bits=nb; c->n_items=n; c->n_cols=nc; return 0; }
int tfim_prep_begin(tfim_ctx *c, const int32_t *ids, int64_t n, const int64_t *off, int64_t rows, int32_t mx, int32_t *freq) {
  c->p_ids.assign(ids, ids + n); c->p_off.assign(off, off + rows + 1); c->p_max = mx;
  memset(freq, 0, (size_t)(mx + 1) * 4);
  for (int64_t i = 0; i < n; i++) { if (ids[i] < 0 || ids[i] > mx) return -2; freq[ids[i]]++; }
  return 0; }
int tfim_prep_finish(tfim_ctx *c, const int32_t *dof, int32_t n_items, int32_t n_cols, int64_t *red, int64_t *lmin,
                     int64_t *lmax, int64_t *lsum, int32_t *dup) {
  std::vector<int64_t> kept; int64_t mn = 0, mx = 0, sm = 0; int d = 0;
  for (int64_t t = 0; t + 1 < (int64_t)c->p_off.size(); t++) {
    int64_t k = 0; for (int64_t j = c->p_off[t]; j < c->p_off[t+1]; j++) k += dof[c->p_ids[j]] >= 0;
    if (k) { mn = kept.empty() ? k : std::min(mn, k); mx = std::max(mx, k); sm += k; kept.push_back(t); } }
  *red = (int64_t)kept.size(); if (lmin) *lmin = mn; if (lmax) *lmax = mx; if (lsum) *lsum = sm;
  if (!kept.empty() && n_items > 0) {
    c->W = ((int64_t)kept.size() + 63) / 64; c->bm.assign((size_t)n_items * c->W, 0);
    for (size_t r = 0; r < kept.size(); r++)
      for (int64_t j = c->p_off[kept[r]]; j < c->p_off[kept[r]+1]; j++) { int x = dof[c->p_ids[j]]; if (x < 0) continue;
        uint64_t &w = c->bm[(size_t)x * c->W + (r >> 6)]; if (w >> (r & 63) & 1) d = 1; w |= 1ull << (r & 63); }
    c->n_bits = (int64_t)kept.size(); c->n_items = n_items; c->n_cols = n_cols; }
  if (dup) *dup = d; c->p_ids.clear(); c->p_off.clear(); return 0; }
int tfim_load_bitmaps_replicated(tfim_ctx *c, const uint64_t *rows, int32_t n, int64_t nb, int64_t w, int32_t nc, int32_t) { return tfim_load_bitmaps(c, rows, n, nb, w, nc); }
static void cnt(tfim_ctx *c, const int32_t *it, int k, const int32_t *cols, int nc, int32_t *out) {
  std::vector<uint64_t> m(c->bm.begin()+(size_t)it[0]*c->W, c->bm.begin()+(size_t)(it[0]+1)*c->W);
  for (int i=1;i<k;i++) for (int64_t w=0;w<c->W;w++) m[w]&=c->bm[(size_t)it[i]*c->W+w];
  for (int j=0;j<nc;j++){ int col=cols?cols[j]:j; const uint64_t*r=&c->bm[(size_t)col*c->W]; int64_t s=0; for(int64_t w=0;w<c->W;w++) s+=__builtin_popcountll(m[w]&r[w]); out[j]=(int32_t)s; } }
int tfim_count_node(tfim_ctx *c, const int32_t *it, int32_t k, int32_t *out) { cnt(c,it,k,nullptr,c->n_cols,out); return 0; }
int tfim_count_nodes(tfim_ctx *c, const int32_t *ids, const int32_t *io, int32_t n, const int32_t *cids, const int32_t *co, int32_t *out, int32_t) {
  for (int r=0;r<n;r++){ if(cids) cnt(c,ids+io[r],io[r+1]-io[r],cids+co[r],co[r+1]-co[r],out+co[r]); else cnt(c,ids+io[r],io[r+1]-io[r],nullptr,c->n_cols,out+(size_t)r*c->n_cols);} return 0; }
static std::vector<int32_t> g_pending[3]; static int64_t g_next=0;
int tfim_count_nodes_submit(tfim_ctx *c, const int32_t *ids, const int32_t *io, int32_t n, const int32_t *cids, const int32_t *co, int32_t p, int64_t *t) {
  int64_t tk=g_next++; auto &v=g_pending[tk%3]; v.assign(cids?co[n]:(size_t)n*c->n_cols,0); tfim_count_nodes(c,ids,io,n,cids,co,v.data(),p); *t=tk; return 0; }
int tfim_count_nodes_ex(tfim_ctx *c, const int32_t *ids, const int32_t *io, int32_t n, const int32_t *cids, const int32_t *co,
                        const int32_t *cr, int32_t nc, const uint64_t *ex, int32_t mc, int32_t *out, int32_t) {
  for (int r=0;r<n;r++) if (co[r+1]>co[r]) cnt(c,ids+io[r],io[r+1]-io[r],cids+co[r],co[r+1]-co[r],out+co[r]);
  const int64_t W=(c->n_cols+63)/64; std::vector<int32_t> all(c->n_cols);
  for (int j=0;j<nc;j++){ const int r=cr[j]; cnt(c,ids+io[r],io[r+1]-io[r],nullptr,c->n_cols,all.data()); const int x0=ids[io[r]];
    int f=0; for (int e=0;e<c->n_cols;e++) if (e!=x0 && !(ex[j*W+(e>>6)]>>(e&63)&1) && all[e]>=mc) f++;
    out[co[n]+2*j]=all[x0]; out[co[n]+2*j+1]=f; }
  return 0; }
int tfim_count_nodes_ex_submit(tfim_ctx *c, const int32_t *ids, const int32_t *io, int32_t n, const int32_t *cids, const int32_t *co,
                               const int32_t *cr, int32_t nc, const uint64_t *ex, int32_t mc, int32_t p, int64_t *t) {
  int64_t tk=g_next++; auto &v=g_pending[tk%3]; v.assign((size_t)co[n]+2*nc,0); tfim_count_nodes_ex(c,ids,io,n,cids,co,cr,nc,ex,mc,v.data(),p); *t=tk; return 0; }
int tfim_count_masks_submit(tfim_ctx *c, const uint64_t *pm, const uint64_t *cm, int32_t n, int32_t, int64_t *t) {
  if (c->n_items > 64) return -4;
  int64_t tk=g_next++; auto &v=g_pending[tk%3]; v.clear(); std::vector<int32_t> it, cols, o;
  for (int r=0;r<n;r++){ it.clear(); cols.clear(); for(uint64_t b=pm[r];b;b&=b-1) it.push_back(__builtin_ctzll(b)); if(it.empty()) return -2;
    for(uint64_t b=cm[r];b;b&=b-1) cols.push_back(__builtin_ctzll(b)); o.assign(cols.size(),0);
    if(!cols.empty()) cnt(c,it.data(),(int)it.size(),cols.data(),(int)cols.size(),o.data()); v.insert(v.end(),o.begin(),o.end()); }
  *t=tk; return 0; }
int tfim_count_masks(tfim_ctx *c, const uint64_t *pm, const uint64_t *cm, int32_t n, int32_t *out, int32_t p) {
  int64_t t; int rc=tfim_count_masks_submit(c,pm,cm,n,p,&t); if(rc) return rc; auto &v=g_pending[t%3]; memcpy(out,v.data(),v.size()*4); return 0; }
int tfim_count_nodes_wait(tfim_ctx *, int64_t t, int32_t *out) { if(t<0) return 0; auto &v=g_pending[t%3]; memcpy(out,v.data(),v.size()*4); return 0; }
int tfim_wave_stats(tfim_ctx *, int64_t *a, int64_t *b, int64_t *r, double *ms) { if(a)*a=0; if(b)*b=0; if(r)*r=0; if(ms)*ms=0; return 0; }
int tfim_wave_graph_stats(tfim_ctx *, int64_t *g, int64_t *l) { if(g)*g=0; if(l)*l=0; return 0; }
int tfim_wave_path_stats(tfim_ctx *, int64_t *a, int64_t *b, int64_t *m, int64_t *f) { if(a)*a=0; if(b)*b=0; if(m)*m=0; if(f)*f=0; return 0; }
// A stand-in network: logits are a fixed hash of each feature row, so the
// learned policy's verdicts (lookahead gate, tail order) land on arbitrary
// nodes and the search logic around them is exercised; training does nothing.
}
struct tfim_mlp { std::vector<float> q; int32_t n = 0; };
static float mock_logit(const float *x, int h) {
  uint64_t z = 0x9E3779B97F4A7C15ull * (uint64_t)(h + 1);
  for (int i = 0; i < TFIM_MLP_INPUTS; i++) { uint32_t b; memcpy(&b, x + i, 4); z = (z ^ b) * 0xBF58476D1CE4E5B9ull; z ^= z >> 29; }
  return (float)((int64_t)(z >> 40) % 601 - 300) / 100.0f;
}
extern "C" {
int tfim_mlp_create(tfim_ctx *, uint64_t, tfim_mlp **o) { *o = new tfim_mlp(); return 0; }
void tfim_mlp_destroy(tfim_mlp *m) { delete m; }
int tfim_mlp_get_params(tfim_mlp *, float *p) { memset(p, 0, TFIM_MLP_PARAMS * 4); return 0; }
int tfim_mlp_set_params(tfim_mlp *, const float *) { return 0; }
int tfim_mlp_forward(tfim_mlp *, const float *x, int32_t n, float *y, int32_t) {
  for (int32_t i = 0; i < n; i++) for (int h = 0; h < 2; h++) y[i * 2 + h] = mock_logit(x + (size_t)i * TFIM_MLP_INPUTS, h);
  return 0; }
int tfim_mlp_forward_submit(tfim_mlp *m, const float *x, int32_t n, int32_t) {
  m->q.assign(x, x + (size_t)n * TFIM_MLP_INPUTS); m->n = n; return 0; }
int tfim_mlp_forward_wait(tfim_mlp *m, float *y) { return tfim_mlp_forward(m, m->q.data(), m->n, y, 1); }
int tfim_mlp_train_step(tfim_mlp *, const float *, const float *, const float *, int32_t, float, float *l) { if (l) *l = 0; return 0; }
int tfim_mlp_train(tfim_mlp *, const float *, const float *, const float *, int32_t, float, int32_t, float *l) { if (l) *l = 0; return 0; }
}
