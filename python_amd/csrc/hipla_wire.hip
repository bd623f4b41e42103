// Binary ragged-batch wire format (host code only).
//
// The reference has no batch format: its matrices live in the caller's arrays, and its tests
// exchange single matrices as "%e" text (the order n, then the row-major values,
// source/tests/test_matrix.c:52-62,113-125), which keeps 7 significant digits.  This format
// keeps a whole ragged batch bit-exact in the layout every entry point of this library takes:
//
//   offset 0   char    magic[8] = "HIPLARB1"
//          8   uint32  version  = 1
//         12   uint32  flags      bit 0: right-hand sides B present
//         16   int64   ncells
//         24   int64   totA       doubles in A (max over cells of offA + n^2)
//         32   int64   totV       doubles in B (max over cells of offV + n)
//         40   uint64  reserved[3] = 0
//         64   int32   n[ncells], then zero padding to a multiple of 8 bytes
//              int64   offA[ncells]
//              int64   offV[ncells]
//              double  A[totA]    cell c at offA[c], row-major
//              double  B[totV]    cell c at offV[c] (when flags bit 0)
//
// All values little-endian (the byte order of every host this library runs on).  A reader
// checks the magic, the version, every cell's extent and the file size before it copies
// anything out.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hipla.h"
#include "hipla_plan.h"

namespace
{
constexpr char kMagic[8] = { 'H', 'I', 'P', 'L', 'A', 'R', 'B', '1' };
constexpr uint32_t kVersion = 1;

struct Header
{
  char magic[8];
  uint32_t version;
  uint32_t flags;
  int64_t ncells, totA, totV;
  uint64_t reserved[3];
};
static_assert (sizeof (Header) == 64, "wire header is 64 bytes");

int64_t pad8 (int64_t bytes) { return (bytes + 7) & ~int64_t (7); }

struct File
{
  std::FILE *f = nullptr;
  ~File ()
  {
    if (f)
      std::fclose (f);
  }
};

bool extents_ok (int64_t ncells, const int *n, const int64_t *offA, const int64_t *offV, int64_t totA, int64_t totV,
                 const char *who)
{
  for (int64_t c = 0; c < ncells; c++)
  {
    // no sum that can wrap: the offsets are compared with the room left below the totals, and n
    // is held to what hipla_plan_create accepts
    const int64_t k = n[c];
    if (k < 1 || k > 65536 || offA[c] < 0 || offV[c] < 0 || k * k > totA || offA[c] > totA - k * k || k > totV ||
        offV[c] > totV - k)
    {
      hipla_set_detail ("%s: cell %lld (n = %lld) lies outside the batch", who, (long long) c, (long long) k);
      return false;
    }
  }
  return true;
}
} // namespace

extern "C" int hipla_batch_write (const char *path, int ncells, const int *n, const int64_t *offA,
                                  const int64_t *offV, const double *A, const double *B)
{
  if (!path || ncells < 0 || (ncells > 0 && (!n || !offA || !offV || !A)))
  {
    hipla_set_detail ("hipla_batch_write: bad argument");
    return 5;
  }
  Header h{};
  std::memcpy (h.magic, kMagic, 8);
  h.version = kVersion;
  h.flags = B ? 1u : 0u;
  h.ncells = ncells;
  for (int c = 0; c < ncells; c++)
  {
    const int64_t k = n[c];
    if (k < 1 || k > 65536 || offA[c] < 0 || offV[c] < 0 || offA[c] > INT64_MAX - k * k || offV[c] > INT64_MAX - k)
    {
      hipla_set_detail ("hipla_batch_write: cell %d has n = %d or an offset out of range", c, n[c]);
      return 5;
    }
    h.totA = std::max (h.totA, offA[c] + k * k);
    h.totV = std::max (h.totV, offV[c] + k);
  }
  File out;
  out.f = std::fopen (path, "wb");
  if (!out.f)
  {
    hipla_set_detail ("hipla_batch_write: cannot open %s", path);
    return 5;
  }
  const int64_t nbytes = (int64_t) sizeof (int) * ncells;
  std::vector<char> zero (pad8 (nbytes) - nbytes, 0);
  bool ok = std::fwrite (&h, sizeof h, 1, out.f) == 1;
  ok = ok && (ncells == 0 || std::fwrite (n, sizeof (int), ncells, out.f) == (size_t) ncells);
  ok = ok && (zero.empty () || std::fwrite (zero.data (), 1, zero.size (), out.f) == zero.size ());
  ok = ok && (ncells == 0 || std::fwrite (offA, sizeof (int64_t), ncells, out.f) == (size_t) ncells);
  ok = ok && (ncells == 0 || std::fwrite (offV, sizeof (int64_t), ncells, out.f) == (size_t) ncells);
  ok = ok && (h.totA == 0 || std::fwrite (A, sizeof (double), h.totA, out.f) == (size_t) h.totA);
  ok = ok && (!B || h.totV == 0 || std::fwrite (B, sizeof (double), h.totV, out.f) == (size_t) h.totV);
  ok = (std::fclose (out.f) == 0) && ok;
  out.f = nullptr;
  if (!ok)
  {
    hipla_set_detail ("hipla_batch_write: short write to %s", path);
    return 5;
  }
  return 0;
}

// Reads and checks the header and the file size; leaves the file at the end of the header.
static int read_header (std::FILE *f, const char *path, Header &h, const char *who)
{
  if (std::fread (&h, sizeof h, 1, f) != 1 || std::memcmp (h.magic, kMagic, 8) != 0)
  {
    hipla_set_detail ("%s: %s is not a ragged-batch file", who, path);
    return 5;
  }
  // sizes far past any file (2^56 doubles) are refused before the byte count is formed, so it
  // cannot overflow
  if (h.version != kVersion || h.ncells < 0 || h.totA < 0 || h.totV < 0 || h.ncells > INT32_MAX ||
      h.totA > (int64_t (1) << 56) || h.totV > (int64_t (1) << 56))
  {
    hipla_set_detail ("%s: %s has version %u / bad sizes", who, path, h.version);
    return 5;
  }
  const int64_t expect = (int64_t) sizeof h + pad8 ((int64_t) sizeof (int) * h.ncells) + 16 * h.ncells +
                         8 * h.totA + ((h.flags & 1u) ? 8 * h.totV : 0);
  if (std::fseek (f, 0, SEEK_END) != 0 || (int64_t) std::ftell (f) != expect ||
      std::fseek (f, (long) sizeof h, SEEK_SET) != 0)
  {
    hipla_set_detail ("%s: %s is truncated or has trailing bytes", who, path);
    return 5;
  }
  return 0;
}

extern "C" int hipla_batch_info (const char *path, int *ncells, int64_t *totA, int64_t *totV, int *has_b)
{
  if (!path || !ncells || !totA || !totV)
  {
    hipla_set_detail ("hipla_batch_info: bad argument");
    return 5;
  }
  File in;
  in.f = std::fopen (path, "rb");
  if (!in.f)
  {
    hipla_set_detail ("hipla_batch_info: cannot open %s", path);
    return 5;
  }
  Header h;
  if (int rc = read_header (in.f, path, h, "hipla_batch_info"))
    return rc;
  *ncells = (int) h.ncells;
  *totA = h.totA;
  *totV = h.totV;
  if (has_b)
    *has_b = (h.flags & 1u) ? 1 : 0;
  return 0;
}

extern "C" int hipla_batch_read (const char *path, int *n, int64_t *offA, int64_t *offV, double *A, double *B)
{
  if (!path)
  {
    hipla_set_detail ("hipla_batch_read: bad argument");
    return 5;
  }
  File in;
  in.f = std::fopen (path, "rb");
  if (!in.f)
  {
    hipla_set_detail ("hipla_batch_read: cannot open %s", path);
    return 5;
  }
  Header h;
  if (int rc = read_header (in.f, path, h, "hipla_batch_read"))
    return rc;
  const int64_t nc = h.ncells;
  if (nc > 0 && (!n || !offA || !offV || !A))
  {
    hipla_set_detail ("hipla_batch_read: bad argument");
    return 5;
  }
  std::vector<int> nv (nc);
  std::vector<int64_t> oa (nc), ov (nc);
  bool ok = nc == 0 || std::fread (nv.data (), sizeof (int), nc, in.f) == (size_t) nc;
  ok = ok && std::fseek (in.f, (long) (sizeof h + pad8 ((int64_t) sizeof (int) * nc)), SEEK_SET) == 0;
  ok = ok && (nc == 0 || std::fread (oa.data (), sizeof (int64_t), nc, in.f) == (size_t) nc);
  ok = ok && (nc == 0 || std::fread (ov.data (), sizeof (int64_t), nc, in.f) == (size_t) nc);
  if (!ok)
  {
    hipla_set_detail ("hipla_batch_read: short read from %s", path);
    return 5;
  }
  if (!extents_ok (nc, nv.data (), oa.data (), ov.data (), h.totA, h.totV, "hipla_batch_read"))
    return 5;
  ok = h.totA == 0 || std::fread (A, sizeof (double), h.totA, in.f) == (size_t) h.totA;
  if (ok && (h.flags & 1u) && B)
    ok = h.totV == 0 || std::fread (B, sizeof (double), h.totV, in.f) == (size_t) h.totV;
  if (!ok)
  {
    hipla_set_detail ("hipla_batch_read: short read from %s", path);
    return 5;
  }
  if (nc > 0)
  {
    std::memcpy (n, nv.data (), sizeof (int) * nc);
    std::memcpy (offA, oa.data (), sizeof (int64_t) * nc);
    std::memcpy (offV, ov.data (), sizeof (int64_t) * nc);
  }
  return 0;
}
