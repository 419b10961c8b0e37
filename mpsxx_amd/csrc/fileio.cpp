// On-disk block-sparse arrays (format in fileio.hpp) and the sweep drivers' file names.
#include "fileio.hpp"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <sstream>

#include "mpsxx_amd.h"

namespace mpx {

namespace {

constexpr char kQstMagic[9] = "MPSXXQST";
constexpr uint32_t kVersion = 1;

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    uint64_t x = h;
    for (size_t i = 0; i < n; ++i) x = (x ^ b[i]) * 1099511628211ull;
    h = x;
  }
};

Error io_error(const std::string& path, const std::string& what) {
  return Error(MPSXX_ERR_IO, path + ": " + what);
}

class Writer {
 public:
  explicit Writer(const std::string& path) : path_(path), tmp_(path + ".part") {
    f_ = std::fopen(tmp_.c_str(), "wb");
    if (!f_) throw io_error(path, std::string("cannot open for writing: ") + std::strerror(errno));
  }
  ~Writer() {
    if (f_) {
      std::fclose(f_);
      std::remove(tmp_.c_str());
    }
  }
  void put(const void* p, size_t n) {
    if (n == 0) return;
    sum_.add(p, n);
    if (std::fwrite(p, 1, n, f_) != n) throw io_error(path_, std::string("write failed: ") + std::strerror(errno));
  }
  template <typename T> void put(const T& v) { put(&v, sizeof(T)); }
  void commit() {
    const uint64_t c = sum_.h;
    if (std::fwrite(&c, 1, 8, f_) != 8) throw io_error(path_, "write failed");
    const int rc = std::fclose(f_);
    f_ = nullptr;
    if (rc != 0) {
      std::remove(tmp_.c_str());
      throw io_error(path_, "close failed");
    }
    if (std::rename(tmp_.c_str(), path_.c_str()) != 0) {
      std::remove(tmp_.c_str());
      throw io_error(path_, std::string("rename failed: ") + std::strerror(errno));
    }
  }

 private:
  std::string path_, tmp_;
  std::FILE* f_ = nullptr;
  Fnv sum_;
};

class Reader {
 public:
  explicit Reader(const std::string& path) : path_(path) {
    f_ = std::fopen(path.c_str(), "rb");
    if (!f_) throw io_error(path, std::string("cannot open: ") + std::strerror(errno));
    if (std::fseek(f_, 0, SEEK_END) != 0) throw io_error(path, "cannot seek");
    const long n = std::ftell(f_);
    if (n < 0) throw io_error(path, "cannot tell size");
    size_ = (uint64_t)n;
    std::rewind(f_);
  }
  ~Reader() {
    if (f_) std::fclose(f_);
  }
  // bytes left before the trailing checksum
  uint64_t left() const { return size_ >= 8 + pos_ ? size_ - 8 - pos_ : 0; }
  void get(void* p, size_t n) {
    if (n == 0) return;
    if (n > left()) throw io_error(path_, "file is truncated");
    if (std::fread(p, 1, n, f_) != n) throw io_error(path_, "read failed");
    sum_.add(p, n);
    pos_ += n;
  }
  template <typename T> T get() {
    T v;
    get(&v, sizeof(T));
    return v;
  }
  // hash the bytes that are not kept (structure-only reads)
  void skip(uint64_t n) {
    std::vector<unsigned char> buf(1 << 20);
    while (n > 0) {
      const size_t k = (size_t)std::min<uint64_t>(n, buf.size());
      get(buf.data(), k);
      n -= k;
    }
  }
  void finish() {
    if (left() != 0) throw io_error(path_, "unexpected bytes after the data");
    uint64_t c = 0;
    if (size_ < 8 || std::fread(&c, 1, 8, f_) != 8) throw io_error(path_, "file is truncated");
    if (c != sum_.h) throw io_error(path_, "checksum mismatch (corrupt file)");
  }
  const std::string& path() const { return path_; }

 private:
  std::string path_;
  std::FILE* f_ = nullptr;
  uint64_t size_ = 0, pos_ = 0;
  Fnv sum_;
};

void check_magic(Reader& r, const char (&magic)[9]) {
  char m[8];
  r.get(m, 8);
  if (std::memcmp(m, magic, 8) != 0) throw io_error(r.path(), std::string("not a ") + magic + " file");
  const uint32_t v = r.get<uint32_t>();
  if (v != kVersion) throw io_error(r.path(), "unsupported format version " + std::to_string(v));
  if (r.get<uint32_t>() != 0) throw io_error(r.path(), "reserved field is not zero");
}

}  // namespace

void write_qst(const std::string& path, const Layout* L, const double* data) {
  Writer w(path);
  w.put(kQstMagic, 8);
  w.put(kVersion);
  w.put((uint32_t)0);
  if (!L) {
    const int32_t z[4] = {0, 0, 0, 0};
    w.put(z, sizeof z);
    w.commit();
    return;
  }
  w.put((int32_t)L->rank);
  w.put((int32_t)(L->qsym ? 1 : 0));
  w.put((int32_t)L->q.p);
  w.put((int32_t)L->q.s);
  for (int i = 0; i < L->rank; ++i) w.put((int32_t)L->nsec(i));
  for (int i = 0; i < L->rank; ++i)
    for (int k = 0; k < L->nsec(i); ++k) {
      const QN q = L->qsym ? L->qs[i][k] : QN();
      w.put((int32_t)q.p);
      w.put((int32_t)q.s);
    }
  for (int i = 0; i < L->rank; ++i)
    for (int k = 0; k < L->nsec(i); ++k) w.put((int32_t)L->ds[i][k]);
  w.put((int64_t)L->nblocks());
  w.put(L->bidx.data(), L->bidx.size() * sizeof(int32_t));
  w.put((int64_t)L->nelem);
  if (L->nelem > 0) {
    if (!data) throw Error(MPSXX_ERR_ARG, "write_qst: null data");
    w.put(data, (size_t)L->nelem * sizeof(double));
  }
  w.commit();
}

QstFile read_qst(const std::string& path, bool want_data) {
  Reader r(path);
  check_magic(r, kQstMagic);
  QstFile out;
  const int32_t rank = r.get<int32_t>();
  const int32_t qsym = r.get<int32_t>();
  const int32_t qp = r.get<int32_t>(), qs = r.get<int32_t>();
  if (rank < 0 || rank > kMaxRank) throw io_error(path, "rank out of range");
  if (qsym != 0 && qsym != 1) throw io_error(path, "invalid symmetry flag");
  if (rank == 0) {
    r.finish();
    return out;
  }
  std::vector<int32_t> nsec(rank);
  r.get(nsec.data(), rank * sizeof(int32_t));
  uint64_t S = 0;
  double ntag = 1.0;
  for (int i = 0; i < rank; ++i) {
    if (nsec[i] < 0) throw io_error(path, "negative sector count");
    S += (uint64_t)nsec[i];
    ntag *= (double)std::max(nsec[i], 1);
  }
  if (ntag > 4.0e18 || 12 * S > r.left()) throw io_error(path, "sector counts out of range");
  std::vector<int32_t> secq(2 * S), secd(S);
  r.get(secq.data(), secq.size() * sizeof(int32_t));
  r.get(secd.data(), secd.size() * sizeof(int32_t));
  std::vector<std::vector<QN>> qv(rank);
  std::vector<std::vector<int>> dv(rank);
  for (int i = 0, k = 0; i < rank; ++i)
    for (int j = 0; j < nsec[i]; ++j, ++k) {
      if (secd[k] < 0) throw io_error(path, "negative sector dimension");
      dv[i].push_back(secd[k]);
      if (qsym) qv[i].push_back(QN(secq[2 * k], secq[2 * k + 1]));
    }
  Layout L;
  L.set_shape(rank, qsym != 0, QN(qp, qs), qv, dv);
  const int64_t nb = r.get<int64_t>();
  if (nb < 0 || (uint64_t)nb > r.left() / (4u * (uint64_t)rank)) throw io_error(path, "block count out of range");
  std::vector<int32_t> bidx((size_t)nb * rank);
  r.get(bidx.data(), bidx.size() * sizeof(int32_t));
  int64_t prev = -1;
  for (int64_t b = 0; b < nb; ++b) {
    const int32_t* idx = bidx.data() + b * rank;
    for (int i = 0; i < rank; ++i)
      if (idx[i] < 0 || idx[i] >= nsec[i]) throw io_error(path, "block index out of range");
    const int64_t t = L.compute_tag(idx);
    if (t <= prev) throw io_error(path, "blocks are not in ascending order");
    prev = t;
    if (!L.allowed(idx)) throw io_error(path, "block is not allowed by the quantum numbers");
    L.add_block(idx);
  }
  L.finalize();
  const int64_t ne = r.get<int64_t>();
  if (ne != L.nelem) throw io_error(path, "element count does not match the blocks");
  if ((uint64_t)ne != r.left() / 8 || r.left() % 8 != 0) throw io_error(path, "data size does not match the blocks");
  if (want_data) {
    out.data.resize((size_t)ne);
    r.get(out.data.data(), (size_t)ne * sizeof(double));
  } else {
    r.skip((uint64_t)ne * sizeof(double));
  }
  r.finish();
  out.empty = false;
  out.L = std::move(L);
  return out;
}

std::string getfile(const std::string& name, const std::string& prefix, int site, int istate, int jstate) {
  std::ostringstream s;
  s << prefix << "/" << name << "." << site << "." << istate << "." << jstate << ".0.tmp";
  return s.str();
}

void tensor_save(const Tensor& t, const std::string& path, hipStream_t s) {
  if (t.empty) {
    write_qst(path, nullptr, nullptr);
    return;
  }
  std::vector<double> host((size_t)t.L.nelem);
  if (t.L.nelem > 0) {
    hip_check(hipMemcpyAsync(host.data(), t.ptr(), host.size() * sizeof(double), hipMemcpyDeviceToHost, s), "save");
    hip_check(hipStreamSynchronize(s), "save sync");
  }
  write_qst(path, &t.L, host.data());
}

void tensor_load(const std::string& path, Tensor& t, hipStream_t s) {
  QstFile f = read_qst(path, true);
  if (f.empty) {
    hip_check(hipStreamSynchronize(s), "load sync");
    t.data.release();
    t.L = Layout();
    t.empty = true;
    return;
  }
  t.set_layout(f.L, s);
  if (f.L.nelem > 0)
    hip_check(hipMemcpyAsync(t.ptr(), f.data.data(), f.data.size() * sizeof(double), hipMemcpyHostToDevice, s), "load");
  hip_check(hipStreamSynchronize(s), "load sync");
}

void write_record(const std::string& path, const char (&magic)[9], const std::vector<int32_t>& fields) {
  Writer w(path);
  w.put(magic, 8);
  w.put(kVersion);
  w.put((uint32_t)0);
  w.put((int32_t)fields.size());
  w.put(fields.data(), fields.size() * sizeof(int32_t));
  w.commit();
}

std::vector<int32_t> read_record(const std::string& path, const char (&magic)[9], size_t nfields) {
  Reader r(path);
  check_magic(r, magic);
  const int32_t n = r.get<int32_t>();
  if (n < 0 || (size_t)n != nfields) throw io_error(path, "unexpected record size");
  std::vector<int32_t> v(nfields);
  r.get(v.data(), nfields * sizeof(int32_t));
  r.finish();
  return v;
}

}  // namespace mpx
