// GPU topology builder: face keys, one stable radix sort, run heads, one scan.
//
//   1. flag the vertex nodes (local index <= dim of some cell); a scan gives
//      them compact ids, so that a face's sorted vertex ids pack in 64 bits
//   2. key[s] for every slot s = e * nF + lf, value s
//   3. stable sort by key (hipCUB): equal keys keep slot order, so a run's
//      first entry is the face's first appearance; runs of three are an error
//   4. isfirst[s] over slot order, exclusive scan -> face ids numbered by
//      first appearance (the host builder's numbering)
//   5. fill faces / face2cell / cell2face from each run head; face_perm by
//      node matching
#include "topology.h"

#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "hfx_internal.h"

namespace hfx {
namespace {

__global__ void vertex_flag_kernel(int nCells, int nN, int nv, const int* __restrict__ cells, int* __restrict__ flag) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nCells * nv) return;
  const int e = (int)(t / nv), k = (int)(t % nv);
  flag[cells[(size_t)e * nN + k]] = 1;
}

__global__ void face_key_kernel(int nCells, int nN, int nF, int dim, const int* __restrict__ cells,
                                const int* __restrict__ fvert, const int* __restrict__ vid,
                                unsigned long long* __restrict__ keys, int* __restrict__ slots) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= (long long)nCells * nF) return;
  const int e = (int)(s / nF), lf = (int)(s % nF);
  unsigned long long v[3] = {0, 0, 0};
  for (int k = 0; k < dim; ++k) v[k] = (unsigned long long)vid[cells[(size_t)e * nN + fvert[lf * dim + k]]];
  // sort the dim (2 or 3) vertex ids
  if (v[0] > v[1]) { const unsigned long long x = v[0]; v[0] = v[1]; v[1] = x; }
  if (dim == 3) {
    if (v[1] > v[2]) { const unsigned long long x = v[1]; v[1] = v[2]; v[2] = x; }
    if (v[0] > v[1]) { const unsigned long long x = v[0]; v[0] = v[1]; v[1] = x; }
    keys[s] = (v[0] << 42) | (v[1] << 21) | v[2];
  } else {
    keys[s] = (v[0] << 32) | v[1];
  }
  slots[s] = (int)s;
}

__global__ void run_kernel(int n, const unsigned long long* __restrict__ keys, const int* __restrict__ slots,
                           int* __restrict__ isfirst, int* __restrict__ partner, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool head = i == 0 || keys[i] != keys[i - 1];
  const int s = slots[i];
  isfirst[s] = head ? 1 : 0;
  if (!head) return;
  const bool pair = i + 1 < n && keys[i + 1] == keys[i];
  partner[s] = pair ? slots[i + 1] : -1;
  if (pair && i + 2 < n && keys[i + 2] == keys[i]) err[0] = 1;
}

__global__ void fill_kernel(int nSlots, int nN, int nF, int nNf, const int* __restrict__ cells,
                            const int* __restrict__ fnodes, const int* __restrict__ isfirst,
                            const int* __restrict__ fid, const int* __restrict__ partner, int* __restrict__ faces,
                            int* __restrict__ face2cell, int* __restrict__ cell2face) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nSlots || !isfirst[s]) return;
  const int f = fid[s], e = s / nF, lf = s % nF, o = partner[s];
  face2cell[2 * (size_t)f] = e;
  face2cell[2 * (size_t)f + 1] = o >= 0 ? o / nF : -1;
  cell2face[s] = f;
  if (o >= 0) cell2face[o] = f;
  for (int j = 0; j < nNf; ++j) faces[(size_t)f * nNf + j] = cells[(size_t)e * nN + fnodes[lf * nNf + j]];
}

__global__ void perm_kernel(int nSlots, int nN, int nF, int nNf, const int* __restrict__ cells,
                            const int* __restrict__ fnodes, const int* __restrict__ cell2face,
                            const int* __restrict__ faces, int8_t* __restrict__ face_perm, int* __restrict__ err) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nSlots * nNf) return;
  const int s = (int)(t / nNf), j = (int)(t % nNf), e = s / nF, lf = s % nF;
  const int node = cells[(size_t)e * nN + fnodes[lf * nNf + j]];
  const int* fn = faces + (size_t)cell2face[s] * nNf;
  int pos = -1;
  for (int k = 0; k < nNf; ++k)
    if (fn[k] == node) pos = k;
  if (pos < 0) err[1] = 1;
  face_perm[t] = (int8_t)(pos < 0 ? 0 : pos);
}

}  // namespace

void build_topology_gpu(const RefEl& re, Mesh& m, int device) {
  const int d = re.dim, nF = re.nF, nNf = re.nNf(), nN = re.nN;
  m.dim = d;
  m.order = re.order;
  m.nN = nN;
  m.nNf = nNf;
  m.nFpc = nF;
  const int nSlots = m.nCells * nF;
  if ((long long)m.nCells * nF > 0x7fffffffLL) throw std::runtime_error("GPU topology: too many faces");
  HFX_HIP(hipSetDevice(device));
  hipStream_t st;
  HFX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct Guard {
    hipStream_t s;
    ~Guard() { (void)hipStreamDestroy(s); }
  } guard{st};
  const int nt = 256;
  DevBuf<int> dcells, dfvert, dfnodes, flag, vid, slots, slots_s, isfirst, fid, partner, faces, f2c, c2f, err;
  DevBuf<unsigned long long> keys, keys_s;
  DevBuf<int8_t> perm;
  DevBuf<unsigned char> tmp;
  dcells.upload(m.cells, st);
  std::vector<int> fv((size_t)nF * d), fnv((size_t)nF * nNf);
  for (int lf = 0; lf < nF; ++lf) {
    for (int k = 0; k < d; ++k) fv[(size_t)lf * d + k] = re.faceNodes[lf][k];  // face vertices come first
    for (int j = 0; j < nNf; ++j) fnv[(size_t)lf * nNf + j] = re.faceNodes[lf][j];
  }
  dfvert.upload(fv, st);
  dfnodes.upload(fnv, st);
  // 1. compact vertex ids
  flag.alloc(m.nNodes + 1);
  vid.alloc(m.nNodes + 1);
  HFX_HIP(hipMemsetAsync(flag.p, 0, (m.nNodes + 1) * sizeof(int), st));
  hipLaunchKernelGGL(vertex_flag_kernel, dim3(ceil_div((long long)m.nCells * (d + 1), nt)), dim3(nt), 0, st,
                     m.nCells, nN, d + 1, dcells.p, flag.p);
  HFX_HIP(hipGetLastError());
  size_t tb = 0;
  HFX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag.p, vid.p, m.nNodes + 1, st));
  tmp.alloc(std::max<size_t>(tb, 1));
  HFX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flag.p, vid.p, m.nNodes + 1, st));
  int nVert = 0;
  HFX_HIP(hipMemcpyAsync(&nVert, vid.p + m.nNodes, sizeof(int), hipMemcpyDeviceToHost, st));
  HFX_HIP(hipStreamSynchronize(st));
  if (d == 3 && nVert >= (1 << 21)) throw std::runtime_error("GPU topology: 3-D meshes need fewer than 2^21 vertices");
  // 2-3. keys, stable sort
  keys.alloc(nSlots);
  keys_s.alloc(nSlots);
  slots.alloc(nSlots);
  slots_s.alloc(nSlots);
  hipLaunchKernelGGL(face_key_kernel, dim3(ceil_div(nSlots, nt)), dim3(nt), 0, st, m.nCells, nN, nF, d, dcells.p,
                     dfvert.p, vid.p, keys.p, slots.p);
  HFX_HIP(hipGetLastError());
  int bits = 1;
  while (bits < 64 && (1ULL << bits) <= (unsigned long long)std::max(nVert, 1)) ++bits;
  const int end_bit = d == 3 ? 42 + bits : 32 + bits;
  tb = 0;
  HFX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.p, keys_s.p, slots.p, slots_s.p, nSlots, 0, end_bit,
                                             st));
  tmp.alloc(std::max<size_t>(tb, tmp.n));
  HFX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.p, keys_s.p, slots.p, slots_s.p, nSlots, 0, end_bit,
                                             st));
  // 4. run heads, face ids by first appearance
  isfirst.alloc(nSlots + 1);
  fid.alloc(nSlots + 1);
  partner.alloc(nSlots);
  err.alloc(2);
  HFX_HIP(hipMemsetAsync(err.p, 0, 2 * sizeof(int), st));
  HFX_HIP(hipMemsetAsync(isfirst.p + nSlots, 0, sizeof(int), st));
  hipLaunchKernelGGL(run_kernel, dim3(ceil_div(nSlots, nt)), dim3(nt), 0, st, nSlots, keys_s.p, slots_s.p, isfirst.p,
                     partner.p, err.p);
  HFX_HIP(hipGetLastError());
  tb = 0;
  HFX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, isfirst.p, fid.p, nSlots + 1, st));
  tmp.alloc(std::max<size_t>(tb, tmp.n));
  HFX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, isfirst.p, fid.p, nSlots + 1, st));
  int hinfo[3] = {0, 0, 0};
  HFX_HIP(hipMemcpyAsync(&hinfo[0], fid.p + nSlots, sizeof(int), hipMemcpyDeviceToHost, st));
  HFX_HIP(hipMemcpyAsync(&hinfo[1], err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HFX_HIP(hipStreamSynchronize(st));
  if (hinfo[1]) throw std::runtime_error("mesh: face shared by more than two cells");
  const int nFaces = hinfo[0];
  // 5. fill
  faces.alloc((size_t)nFaces * nNf);
  f2c.alloc((size_t)nFaces * 2);
  c2f.alloc(nSlots);
  perm.alloc((size_t)nSlots * nNf);
  hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(nSlots, nt)), dim3(nt), 0, st, nSlots, nN, nF, nNf, dcells.p,
                     dfnodes.p, isfirst.p, fid.p, partner.p, faces.p, f2c.p, c2f.p);
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(perm_kernel, dim3(ceil_div((long long)nSlots * nNf, nt)), dim3(nt), 0, st, nSlots, nN, nF, nNf,
                     dcells.p, dfnodes.p, c2f.p, faces.p, perm.p, err.p);
  HFX_HIP(hipGetLastError());
  m.nFaces = nFaces;
  m.faces.resize((size_t)nFaces * nNf);
  m.face2cell.resize((size_t)nFaces * 2);
  m.cell2face.resize(nSlots);
  m.face_perm.resize((size_t)nSlots * nNf);
  faces.download(m.faces.data(), m.faces.size(), st);
  f2c.download(m.face2cell.data(), m.face2cell.size(), st);
  c2f.download(m.cell2face.data(), m.cell2face.size(), st);
  perm.download(m.face_perm.data(), m.face_perm.size(), st);
  HFX_HIP(hipMemcpyAsync(&hinfo[2], err.p + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  HFX_HIP(hipStreamSynchronize(st));
  if (hinfo[2]) throw std::runtime_error("mesh: cell node not found in face");
  m.boundary.clear();
  for (int f = 0; f < nFaces; ++f)
    if (m.face2cell[2 * (size_t)f + 1] == -1) m.boundary.push_back(f);
}

}  // namespace hfx
