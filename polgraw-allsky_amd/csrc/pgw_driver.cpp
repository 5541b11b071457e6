// The search driver around the batched engine: sky loops, trigger files and checkpoints
// (search(), search/network/src-cpu/jobcore.c:51-184; read_checkpoints(), init.c:868-917).
// Host code only; the engine calls go through the library's own C ABI.
#include <errno.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pgw.h"

namespace {

// Appends nrec records of PGW_NPAR values of `esize` bytes (8: double, 4: float) to `path` under a
// whole-file fcntl write lock (jobcore.c:127-145).
int append_records(const std::string& path, const void* v, size_t nrec, size_t esize) {
  const int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_APPEND, S_IRUSR | S_IWUSR | S_IRGRP | S_IROTH);
  if (fd < 0) return PGW_ERR_IO;
  struct flock lck = {};
  lck.l_type = F_WRLCK;
  lck.l_whence = SEEK_SET;
  lck.l_start = 0;
  lck.l_len = 0;
  int st = PGW_OK;
  while (fcntl(fd, F_SETLKW, &lck) < 0)
    if (errno != EINTR) { st = PGW_ERR_IO; break; }
  const char* p = reinterpret_cast<const char*>(v);
  size_t left = nrec * PGW_NPAR * esize;
  while (!st && left > 0) {
    const ssize_t w = write(fd, p, left);
    if (w < 0) {
      if (errno == EINTR) continue;
      st = PGW_ERR_IO;
      break;
    }
    p += w;
    left -= (size_t)w;
  }
  if (close(fd) < 0 && !st) st = PGW_ERR_IO;
  return st;
}

// "pm mm nn sst FNum" (jobcore.c:100-104), replaced atomically
int write_state(const char* qname, int pm, int mm, int nn, int sst, int FNum) {
  const std::string tmp = std::string(qname) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return PGW_ERR_IO;
  const bool ok = fprintf(f, "%d %d %d %d %d\n", pm, mm, nn, sst, FNum) > 0;
  if (fclose(f) != 0 || !ok) return PGW_ERR_IO;
  return rename(tmp.c_str(), qname) == 0 ? PGW_OK : PGW_ERR_IO;
}

struct Sky {
  int pm, mm, nn;
};

}  // namespace

extern "C" {

int pgw_read_checkpoint(const char* qname, pgw_range* r, int* FNum) {
  if (!qname || !r || !FNum) return PGW_ERR_ARG;
  FILE* f = fopen(qname, "r");
  if (!f) {  // no state file: start from the beginning
    r->pst = r->pmr[0];
    r->mst = r->mr[0];
    r->nst = r->nr[0];
    r->sst = r->spndr[0];
    *FNum = 0;
    return PGW_OK;
  }
  int v[5];
  const int n = fscanf(f, "%d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4]);
  fclose(f);
  if (n == EOF) return 1;  // empty: the run has finished
  if (n != 5) return PGW_ERR_IO;
  r->pst = v[0];
  r->mst = v[1];
  r->nst = v[2];
  r->sst = v[3];
  *FNum = v[4];
  return PGW_OK;
}

int pgw_search(pgw_handle* h, pgw_range* r, const pgw_opts* opts, const pgw_output* out, int* FNum) {
  if (!h || !r || !opts || !out || !out->prefix || !FNum) return PGW_ERR_ARG;
  pgw_settings S;
  int st = pgw_get_settings(h, 0, &S);
  if (st) return st;
  const int chunk = out->checkpoint_every > 0 ? out->checkpoint_every : 64;
  const char* label = out->label ? out->label : "";
  // records as doubles, or as floats for a FLOAT_TYPE = float driver (search/network/src-cpu/auxi.h:16-23)
  const bool fl = out->float_records != 0;
  const size_t es = fl ? sizeof(float) : sizeof(double);
  std::vector<char> buf;                          // records not yet in the trigger file (bytes)
  std::vector<char> part((size_t)PGW_NPAR * 2 * S.nfft * es);
  int done = 0;
  for (int pm = r->pst; pm <= r->pmr[1]; ++pm) {
    char name[1024];
    snprintf(name, sizeof name, "%s/triggers_%03d_%04d%s_%d.bin", out->prefix, out->ident, out->band, label, pm);
    const std::string path(name);
    // this hemisphere's sky points in the reference's loop order from the current starts
    std::vector<Sky> sky;
    for (int mm = r->mst; mm <= r->mr[1]; ++mm) {
      for (int nn = r->nst; nn <= r->nr[1]; ++nn) sky.push_back(Sky{pm, mm, nn});
      r->nst = r->nr[0];
    }
    r->mst = r->mr[0];
    size_t k = 0;
    while (k < sky.size()) {
      if (out->max_sky > 0 && done >= out->max_sky)  // the state (if any) already points at sky[k]
        return append_records(path, buf.data(), buf.size() / (PGW_NPAR * es), es);
      // the first sky point after a resume keeps its own first spindown (sst), the rest start
      // at spndr[0], as search() resets sst after each job_core
      size_t n = r->sst != r->spndr[0] ? 1 : std::min(sky.size() - k, (size_t)chunk);
      if (out->max_sky > 0) n = std::min(n, (size_t)(out->max_sky - done));
      const int FN0 = *FNum;
      int cnt = 0;
      for (;;) {
        const int room = (int)(part.size() / (PGW_NPAR * es));
        st = fl ? pgw_job_batch_f(h, &sky[k].pm, (int)n, r, opts, &cnt, reinterpret_cast<float*>(part.data()), room, FNum)
                : pgw_job_batch(h, &sky[k].pm, (int)n, r, opts, &cnt, reinterpret_cast<double*>(part.data()), room, FNum);
        if (st != PGW_ERR_CAPACITY) break;
        if (cnt <= room) return PGW_ERR_CAPACITY;  // no larger count to make room for: do not spin
        part.resize((size_t)PGW_NPAR * cnt * es);  // room for every record of the batch, then again
        cnt = 0;
        *FNum = FN0;
      }
      if (st) return st;
      r->sst = r->spndr[0];
      buf.insert(buf.end(), part.begin(), part.begin() + (size_t)PGW_NPAR * cnt * es);
      k += n;
      done += (int)n;
      const bool last = k == sky.size();
      // flush rule of jobcore.c:127 (more than nfft buffered), and before every state update
      if (buf.size() / (PGW_NPAR * es) > (size_t)S.nfft || (out->qname && !last)) {
        if ((st = append_records(path, buf.data(), buf.size() / (PGW_NPAR * es), es))) return st;
        buf.clear();
      }
      if (out->qname && !last && (st = write_state(out->qname, pm, sky[k].mm, sky[k].nn, r->sst, *FNum)))
        return st;
    }
    // leftover of the hemisphere (jobcore.c:152-166); creates the file even when empty
    if ((st = append_records(path, buf.data(), buf.size() / (PGW_NPAR * es), es))) return st;
    buf.clear();
    r->pst = pm + 1;
    if (out->qname && pm < r->pmr[1] && (st = write_state(out->qname, pm + 1, r->mr[0], r->nr[0], r->sst, *FNum)))
      return st;
  }
  // finished: the state file is emptied (main.c:141-147)
  if (out->qname) {
    FILE* f = fopen(out->qname, "w");
    if (!f || fclose(f) != 0) return PGW_ERR_IO;
  }
  return PGW_OK;
}

}  // extern "C"
