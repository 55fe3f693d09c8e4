// fr_output.cpp -- solution files and restart.
//
//  * ParaView .vtu (writeParaview, output.cpp:103-391): one <Piece> per
//    element with the primitives (density, velocity, pressure) at the
//    (P+3)^3 plot points -- the tensor grid of {-1, spts, +1} -- interpolated
//    from the solution points (extrapolateUPpts, solver.cpp:449-484), their
//    positions (shape functions of the linear element, solver.cpp:904-951)
//    and the (P+2)^3 linear sub-hexes.  3D only, as there.
//  * CSV (writeCSV, output.cpp:26-100): position and primitives at every
//    solution point, then at every flux point.
//  * Restart from a .vtu (solver::readRestartFile, solver.cpp:1420-1482;
//    ele::restart, ele.cpp:1451-1880): the primitives at the solution points
//    are picked out of the plot points, converted to conservative variables
//    and, when the file's order differs, interpolated to this order's points.
//  * An exact binary checkpoint of U_spts (no reference counterpart).
//
// The FP64 state is copied from the device; the formatting, the primitives
// and the small tensor interpolations run on the host threads, one element
// batch at a time.  Numbers are written as the reference's streams write
// them: precision(16) general format in the .vtu ("%.16g"), precision(15)
// fixed in the CSV ("%.15f").
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "fr_ctx.hpp"

using namespace frimpl;

namespace {

// Host copy of the FP64 state in mesh element order: U[e][field][spt].
int hostState(fr_ctx *c, std::vector<double> &Ue) {
  const size_t per = (size_t)c->NF * c->NS;
  std::vector<double> h(per * c->nE);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(h.data(), c->U, h.size() * 8, hipMemcpyDeviceToHost));
  Ue.resize(h.size());
  for (int e = 0; e < c->nE; e++) std::memcpy(&Ue[per * e], &h[per * c->iperm[e]], per * 8);
  return 0;
}

// Node coordinates of mesh element e: X[n][3].
const double *nodesOf(fr_ctx *c, int e) { return &c->X[(size_t)c->iperm[e] * c->ops.nNodes * 3]; }

// Primitives (ele::getPrimitives, ele.cpp:1107-1126) of the conservative
// state u[k] (stride `st` between fields) into v[k].
void primitives(const fr_ctx *c, const double *u, size_t st, double *v) {
  const int D = c->D;
  if (c->cfg.equation == 0) {
    v[0] = u[0];
    return;
  }
  const double rho = u[0];
  double vsq = 0.0;
  for (int d = 0; d < D; d++) {
    v[1 + d] = u[(1 + d) * st] / rho;
    vsq += v[1 + d] * v[1 + d];
  }
  v[0] = rho;
  v[D + 1] = (c->cfg.gamma - 1.0) * (u[(D + 1) * st] - 0.5 * rho * vsq);
}

// 1D interpolation matrix from the nodes `from` to the points `to`:
// A[a][i] = l_i(to[a]).
std::vector<double> interp1D(const std::vector<double> &from, const std::vector<double> &to) {
  std::vector<double> A(to.size() * from.size());
  for (size_t a = 0; a < to.size(); a++)
    for (size_t i = 0; i < from.size(); i++) A[a * from.size() + i] = frhost::lagrange(from, to[a], (int)i);
  return A;
}

// Tensor-product interpolation of one 3D field: in[(k*n + j)*n + i] on n^3
// nodes -> out[(c*m + b)*m + a] on m^3 points, with A (m x n), as the dense
// product with opp_spts_to_ppts (operators.cpp:111-143) factorised by axis.
void tensor3(const double *A, int m, int n, const double *in, size_t inStride, double *out, size_t outStride) {
  std::vector<double> t1((size_t)m * n * n), t2((size_t)m * m * n);
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int a = 0; a < m; a++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += A[a * n + i] * in[((size_t)(k * n + j) * n + i) * inStride];
        t1[((size_t)k * n + j) * m + a] = s;
      }
  for (int k = 0; k < n; k++)
    for (int b = 0; b < m; b++)
      for (int a = 0; a < m; a++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += A[b * n + j] * t1[((size_t)k * n + j) * m + a];
        t2[((size_t)k * m + b) * m + a] = s;
      }
  for (int cc = 0; cc < m; cc++)
    for (int b = 0; b < m; b++)
      for (int a = 0; a < m; a++) {
        double s = 0.0;
        for (int k = 0; k < n; k++) s += A[cc * n + k] * t2[((size_t)k * m + b) * m + a];
        out[((size_t)(cc * m + b) * m + a) * outStride] = s;
      }
}

void putG16(std::string &s, double v) {
  char b[40];
  const int n = std::snprintf(b, sizeof b, "%.16g ", v);
  s.append(b, n);
}
void putF15(std::string &s, double v, char sep) {
  char b[64];
  const int n = std::snprintf(b, sizeof b, "%.15f%c", v, sep);
  s.append(b, n);
}
void putI(std::string &s, long v) {
  char b[24];
  const int n = std::snprintf(b, sizeof b, "%ld ", v);
  s.append(b, n);
}

// Formats elements [e0, e1) with f(e, string&) on up to 16 threads and
// appends the pieces to `out` in element order.
template <class F>
void formatElements(int e0, int e1, F f, std::ofstream &out) {
  const int nT = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  const int batch = 2048;
  std::vector<std::string> part(nT);
  for (int b0 = e0; b0 < e1; b0 += batch * nT) {
    std::vector<std::thread> th;
    for (int t = 0; t < nT; t++) {
      const int lo = b0 + t * batch, hi = std::min(e1, lo + batch);
      part[t].clear();
      if (lo >= hi) continue;
      th.emplace_back([&, t, lo, hi] {
        for (int e = lo; e < hi; e++) f(e, part[t]);
      });
    }
    for (auto &x : th) x.join();
    for (int t = 0; t < nT; t++) out.write(part[t].data(), (std::streamsize)part[t].size());
  }
}

std::string iterName(const char *base, const char *fmt, int iter) {
  std::vector<char> b(std::strlen(base) + 64);
  std::snprintf(b.data(), b.size(), fmt, base, iter);
  return std::string(b.data());
}

int writeParaview(fr_ctx *c, const char *name, int iter, double time) {
  if (c->D != 3) return fail(FR_EINVAL, "Invalid dimensionality [nDims].");
  std::vector<double> Ue;
  if (int rc = hostState(c, Ue)) return rc;
  const int P = c->P, P1 = P + 1, NP = P + 3, nPpts = NP * NP * NP, nSub = (P + 2) * (P + 2) * (P + 2);
  const int NF = c->NF, NS = c->NS, nN = c->ops.nNodes;
  std::vector<double> p1d(NP);
  p1d[0] = -1.0;
  for (int i = 0; i < P1; i++) p1d[1 + i] = c->ops.spts1D[i];
  p1d[NP - 1] = 1.0;
  const std::vector<double> A = interp1D(c->ops.spts1D, p1d);
  std::vector<double> shp((size_t)nPpts * nN);  // shape_ppts(ppt, node)
  for (int q = 0; q < nPpts; q++) {
    const double r[3] = {p1d[q % NP], p1d[(q / NP) % NP], p1d[q / (NP * NP)]};
    frhost::shapeFunctions(3, r, &shp[(size_t)q * nN], nN);
  }
  // sub-cell connectivity, offsets and types: the same for every element
  std::string cells;
  for (int k = 0; k < NP - 1; k++)
    for (int j = 0; j < NP - 1; j++) {
      for (int i = 0; i < NP - 1; i++) {
        const long q[8] = {i + NP * (j + NP * k),           i + 1 + NP * (j + NP * k),
                           i + 1 + NP * (j + 1 + NP * k),   i + NP * (j + 1 + NP * k),
                           i + NP * (j + NP * (k + 1)),     i + 1 + NP * (j + NP * (k + 1)),
                           i + 1 + NP * (j + 1 + NP * (k + 1)), i + NP * (j + 1 + NP * (k + 1))};
        for (long v : q) putI(cells, v);
        cells += "\n";
      }
    }
  std::string offs, types;
  for (int k = 0; k < nSub; k++) {
    putI(offs, (long)(k + 1) * 8);
    types += "12 ";
  }
  const bool ns = c->cfg.equation != 0;

  std::ofstream out(iterName(name, "%s_%.09d.vtu", iter), std::ios::binary);
  if (!out) return fail(FR_EIO, "cannot open the .vtu file for writing");
  std::string head = "<?xml version=\"1.0\" ?>\n"
                     "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
                     "compressor=\"vtkZLibDataCompressor\">\n";
  char tb[64];
  std::snprintf(tb, sizeof tb, "%.16g", time);
  head += std::string("<!-- TIME ") + tb + " -->\n";
  head += "<!-- ITER " + std::to_string(iter) + " -->\n";
  head += "\t<UnstructuredGrid>\n";
  out << head;

  auto piece = [&](int e, std::string &s) {
    const double *u = &Ue[(size_t)e * NF * NS];
    std::vector<double> V((size_t)NF * NS), Vp((size_t)NF * nPpts);
    std::vector<double> tmp(NF);
    for (int sp = 0; sp < NS; sp++) {
      primitives(c, u + sp, NS, tmp.data());
      for (int k = 0; k < NF; k++) V[(size_t)k * NS + sp] = tmp[k];
    }
    for (int k = 0; k < NF; k++) tensor3(A.data(), NP, P1, &V[(size_t)k * NS], 1, &Vp[(size_t)k * nPpts], 1);
    s += "\t\t<Piece NumberOfPoints=\"" + std::to_string(nPpts) + "\" NumberOfCells=\"" + std::to_string(nSub) +
         "\">\n";
    s += "\t\t\t<PointData>\n";
    s += "\t\t\t\t<DataArray type=\"Float32\" Name=\"Density\" format=\"ascii\">\n";
    for (int q = 0; q < nPpts; q++) putG16(s, Vp[q]);
    s += "\n\t\t\t\t</DataArray>\n";
    if (ns) {
      s += "\t\t\t\t<DataArray type=\"Float32\" NumberOfComponents=\"3\" Name=\"Velocity\" format=\"ascii\">\n";
      for (int q = 0; q < nPpts; q++)
        for (int d = 1; d <= 3; d++) putG16(s, Vp[(size_t)d * nPpts + q]);
      s += "\n\t\t\t\t</DataArray>\n";
      s += "\t\t\t\t<DataArray type=\"Float32\" Name=\"Pressure\" format=\"ascii\">\n";
      for (int q = 0; q < nPpts; q++) putG16(s, Vp[(size_t)4 * nPpts + q]);
      s += "\n\t\t\t\t</DataArray>\n";
    }
    s += "\t\t\t</PointData>\n";
    s += "\t\t\t<Points>\n";
    s += "\t\t\t\t<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    const double *X = nodesOf(c, e);
    for (int q = 0; q < nPpts; q++)
      for (int d = 0; d < 3; d++) {
        double x = 0.0;
        for (int n = 0; n < nN; n++) x += shp[(size_t)q * nN + n] * X[n * 3 + d];
        putG16(s, x);
      }
    s += "\n\t\t\t\t</DataArray>\n";
    s += "\t\t\t</Points>\n";
    s += "\t\t\t<Cells>\n";
    s += "\t\t\t\t<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    s += cells;
    s += "\t\t\t\t</DataArray>\n";
    s += "\t\t\t\t<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    s += offs;
    s += "\n\t\t\t\t</DataArray>\n";
    s += "\t\t\t\t<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    s += types;
    s += "\n\t\t\t\t</DataArray>\n";
    s += "\t\t\t</Cells>\n";
    s += "\t\t</Piece>\n";
  };
  formatElements(0, c->nE, piece, out);
  out << "\t</UnstructuredGrid>\n</VTKFile>\n";
  out.close();
  if (!out) return fail(FR_EIO, "writing the .vtu file failed");
  return 0;
}

int writeCSV(fr_ctx *c, const char *name, int iter) {
  std::vector<double> Ue;
  if (int rc = hostState(c, Ue)) return rc;
  const int D = c->D, NF = c->NF, NS = c->NS, NFP = c->NFP, nN = c->ops.nNodes;
  std::vector<double> shS((size_t)NS * nN), shF((size_t)NFP * nN);
  for (int p = 0; p < NS; p++) frhost::shapeFunctions(D, &c->ops.locSpts[3 * p], &shS[(size_t)p * nN], nN);
  for (int p = 0; p < NFP; p++) frhost::shapeFunctions(D, &c->ops.locFpts[3 * p], &shF[(size_t)p * nN], nN);
  const std::vector<double> &E = c->ops.spts_to_fpts;  // nFpts x nSpts

  std::ofstream out(iterName(name, "%s.csv.%.09d", iter), std::ios::binary);
  if (!out) return fail(FR_EIO, "cannot open the .csv file for writing");
  out << "x,y,z,";
  out << (c->cfg.equation == 0 ? "rho\n" : "rho,u,v,p\n");

  auto row = [&](std::string &s, const double *sh, const double *X, const double *v) {
    for (int d = 0; d < D; d++) {
      double x = 0.0;
      for (int n = 0; n < nN; n++) x += sh[n] * X[n * 3 + d];
      putF15(s, x, ',');
    }
    if (D == 2) s += "0.0,";
    for (int k = 0; k < NF; k++) putF15(s, v[k], k + 1 < NF ? ',' : '\n');
  };
  auto rows = [&](int e, std::string &s) {
    const double *u = &Ue[(size_t)e * NF * NS];
    const double *X = nodesOf(c, e);
    std::vector<double> v(NF), uf(NF);
    for (int sp = 0; sp < NS; sp++) {
      primitives(c, u + sp, NS, v.data());
      row(s, &shS[(size_t)sp * nN], X, v.data());
    }
    // U_fpts = opp_spts_to_fpts U_spts (solver::extrapolateU, solver.cpp:401-412)
    for (int q = 0; q < NFP; q++) {
      for (int k = 0; k < NF; k++) {
        double acc = 0.0;
        for (int sp = 0; sp < NS; sp++) acc += E[(size_t)q * NS + sp] * u[(size_t)k * NS + sp];
        uf[k] = acc;
      }
      primitives(c, uf.data(), 1, v.data());
      row(s, &shF[(size_t)q * nN], X, v.data());
    }
  };
  formatElements(0, c->nE, rows, out);
  out.close();
  if (!out) return fail(FR_EIO, "writing the .csv file failed");
  return 0;
}

// Value of attribute `key` ("NumberOfCells") in a tag line, or -1.
long attrValue(const std::string &line, const char *key) {
  const std::string k = std::string(key) + "=\"";
  const size_t p = line.find(k);
  if (p == std::string::npos) return -1;
  return std::strtol(line.c_str() + p + k.size(), nullptr, 10);
}

// Name="..." of a <DataArray line, or "".
std::string arrayName(const std::string &line) {
  const size_t p = line.find("Name=\"");
  if (p == std::string::npos) return "";
  const size_t q = line.find('"', p + 6);
  return line.substr(p + 6, q == std::string::npos ? std::string::npos : q - p - 6);
}

int readRestart(fr_ctx *c, const char *name, int restartIter, double *time) {
  if (c->D != 3) return fail(FR_EINVAL, "restart files are 3D (.vtu written by the 3D writer)");
  std::ifstream in(iterName(name, "%s_%.09d.vtu", restartIter));
  if (!in) return fail(FR_EIO, "Cannot open restart file.");
  std::string line;
  bool foundUG = false;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string w;
    ss >> w;
    if (w == "<!--") {
      ss >> w;
      double t;
      if (w == "TIME" && (ss >> t) && time) *time = t;
    } else if (w == "<UnstructuredGrid>") {
      foundUG = true;
      break;
    }
  }
  if (!foundUG) return fail(FR_EIO, "Cannot find UnstructuredData tag in restart file.");

  const int P = c->P, P1 = P + 1, NS = c->NS, NF = c->NF, nE = c->nE;
  const bool ns = c->cfg.equation != 0;
  const double gm1 = c->cfg.gamma - 1.0;
  std::vector<double> Uref((size_t)NS * nE * NF);  // reference layout [spt][ele][field]
  std::vector<std::vector<double>> interpCache(8);
  for (int e = 0; e < nE; e++) {
    if (!std::getline(in, line)) return fail(FR_EIO, "Restart-file element doesn't exist!");
    const long nCells = attrValue(line, "NumberOfCells");
    if (nCells <= 0) return fail(FR_EIO, "Restart-file element doesn't exist!");
    const int Pr = (int)std::lround(std::cbrt((double)nCells)) - 2, Pr1 = Pr + 1, NPr = Pr + 3;
    if (Pr < 0 || Pr > 6 || (long)(Pr + 2) * (Pr + 2) * (Pr + 2) != nCells)
      return fail(FR_EIO, "restart element has an unsupported number of plot cells");
    const int nSr = Pr1 * Pr1 * Pr1;
    // tempU[spt][field] at the file's solution points (interior plot points)
    std::vector<double> tu((size_t)nSr * NF, 0.0), tv((size_t)nSr * 3, 0.0), tp(nSr, 0.0);
    bool fRho = false, fV = false, fP = false;
    const size_t nPr = (size_t)NPr * NPr * NPr;
    auto interior = [&](int s) {
      const int i = s % Pr1, j = (s / Pr1) % Pr1, k = s / (Pr1 * Pr1);
      return (size_t)(i + 1) + NPr * ((j + 1) + (size_t)NPr * (k + 1));
    };
    while (!(fRho && (!ns || (fV && fP)))) {
      std::string nm;
      while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string w;
        ss >> w;
        if (w == "<DataArray") {
          nm = arrayName(line);
          break;
        }
        if (w == "</Piece>") return fail(FR_EIO, "restart element lacks density, velocity or pressure");
      }
      if (nm.empty()) return fail(FR_EIO, "restart file ends inside an element");
      if (!std::getline(in, line)) return fail(FR_EIO, "restart file ends inside a data array");
      const int ncomp = (nm == "Velocity") ? 3 : 1;
      if (nm != "Density" && nm != "Velocity" && nm != "Pressure") continue;
      std::vector<double> vals(nPr * ncomp);
      const char *p = line.c_str();
      for (size_t i = 0; i < vals.size(); i++) {
        char *endp;
        vals[i] = std::strtod(p, &endp);
        if (endp == p) return fail(FR_EIO, "short data array in restart file");
        p = endp;
      }
      for (int s = 0; s < nSr; s++) {
        const size_t q = interior(s);
        if (nm == "Density") tu[(size_t)s * NF] = vals[q];
        else if (nm == "Pressure") tp[s] = vals[q];
        else
          for (int d = 0; d < 3; d++) tv[(size_t)s * 3 + d] = vals[q * 3 + d];
      }
      if (nm == "Density") fRho = true;
      else if (nm == "Velocity") fV = true;
      else fP = true;
      if (!ns && fRho) break;
    }
    // to the end of this element's data
    while (line.find("</Piece>") == std::string::npos)
      if (!std::getline(in, line)) return fail(FR_EIO, "restart file ends inside an element");
    if (ns)
      for (int s = 0; s < nSr; s++) {
        const double rho = tu[(size_t)s * NF];
        double vsq = 0.0;
        for (int d = 0; d < 3; d++) {
          tu[(size_t)s * NF + 1 + d] = rho * tv[(size_t)s * 3 + d];
          vsq += tv[(size_t)s * 3 + d] * tv[(size_t)s * 3 + d];
        }
        tu[(size_t)s * NF + 4] = tp[s] / gm1 + 0.5 * rho * vsq;
      }
    // restart order -> this order (opp_interp, ele.cpp:1505-1557)
    std::vector<double> fin((size_t)NS * NF);
    if (Pr == P) {
      fin = tu;
    } else {
      auto &Ai = interpCache[Pr];
      if (Ai.empty()) Ai = interp1D(frhost::pts1D(c->cfg.sptsType, Pr), c->ops.spts1D);
      for (int k = 0; k < NF; k++) tensor3(Ai.data(), P1, Pr1, &tu[k], NF, &fin[k], NF);
    }
    for (int s = 0; s < NS; s++)
      for (int k = 0; k < NF; k++) Uref[((size_t)s * nE + e) * NF + k] = fin[(size_t)s * NF + k];
  }
  return fr_set_state(c, Uref.data(), FR_LAYOUT_REF);
}

const char kMagic[8] = {'F', 'R', 'C', 'K', 'P', 'T', '0', '1'};

struct CkptHeader {
  char magic[8];
  int32_t nDims, order, nFields, nEles, nSpts, iter;
  double time;
};

}  // namespace

extern "C" {

int fr_write_data(fr_ctx *c, const char *dataFileName, int plotType, int iter, double time) {
  if (!c || !dataFileName) return fail(FR_EINVAL, "null argument");
  if (plotType == 0) return writeCSV(c, dataFileName, iter);
  if (plotType == 1) return writeParaview(c, dataFileName, iter, time);
  return fail(FR_EINVAL, "plotType must be 0 (CSV) or 1 (ParaView)");
}

int fr_read_restart(fr_ctx *c, const char *dataFileName, int restartIter, double *time) {
  if (!c || !dataFileName) return fail(FR_EINVAL, "null argument");
  return readRestart(c, dataFileName, restartIter, time);
}

int fr_write_checkpoint(fr_ctx *c, const char *path, int iter, double time) {
  if (!c || !path) return fail(FR_EINVAL, "null argument");
  std::vector<double> U(c->nState());
  if (int rc = fr_get_state(c, U.data(), FR_LAYOUT_REF)) return rc;
  CkptHeader h{};
  std::memcpy(h.magic, kMagic, 8);
  h.nDims = c->D;
  h.order = c->P;
  h.nFields = c->NF;
  h.nEles = c->nE;
  h.nSpts = c->NS;
  h.iter = iter;
  h.time = time;
  std::ofstream out(path, std::ios::binary);
  if (!out) return fail(FR_EIO, "cannot open the checkpoint file for writing");
  out.write((const char *)&h, sizeof h);
  out.write((const char *)U.data(), (std::streamsize)(U.size() * 8));
  out.close();
  if (!out) return fail(FR_EIO, "writing the checkpoint file failed");
  return 0;
}

int fr_read_checkpoint(fr_ctx *c, const char *path, int *iter, double *time) {
  if (!c || !path) return fail(FR_EINVAL, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(FR_EIO, "cannot open the checkpoint file");
  CkptHeader h{};
  in.read((char *)&h, sizeof h);
  if (!in || std::memcmp(h.magic, kMagic, 8) != 0) return fail(FR_EIO, "not a checkpoint file");
  if (h.nDims != c->D || h.order != c->P || h.nFields != c->NF || h.nEles != c->nE || h.nSpts != c->NS)
    return fail(FR_EINVAL, "checkpoint does not match this mesh / order / equation");
  std::vector<double> U(c->nState());
  in.read((char *)U.data(), (std::streamsize)(U.size() * 8));
  if (!in) return fail(FR_EIO, "checkpoint file is truncated");
  if (iter) *iter = h.iter;
  if (time) *time = h.time;
  return fr_set_state(c, U.data(), FR_LAYOUT_REF);
}

}  // extern "C"
