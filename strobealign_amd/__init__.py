"""MI355X (gfx950) drop-in for strobealign's read-mapping path.

The mapping path -- randstrobe seeding, index lookups, hit filtering and
rescue, anchor chaining into NAMs, and SSW / piecewise extension -- runs as HIP
kernels in ``lib/libsa_hip.so``; the host library replays the reference's
per-chunk decisions and writes SAM, PAF or abundance output. This module is a
thin ctypes binding of the C ABI in ``include/sa_hip.h``; it raises instead of
falling back when the library is missing or no GPU is present.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsa_hip.so")
CLI_PATH = os.path.join(HERE, "bin", "strobealign-hip")
CSRC = os.path.join(HERE, "csrc")

_lib = None


class SaError(RuntimeError):
    pass


class SeedingParams(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int32), ("s", ctypes.c_int32), ("t", ctypes.c_int32), ("w_min", ctypes.c_int32),
                ("w_max", ctypes.c_int32), ("max_dist", ctypes.c_int32), ("q", ctypes.c_uint64),
                ("main_hash_mask", ctypes.c_uint64), ("forward_main_hash_mask", ctypes.c_uint64),
                ("partial_orientation_pos", ctypes.c_int32), ("bits", ctypes.c_int32),
                ("filter_cutoff", ctypes.c_uint64)]


class MapOptions(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("eqx", ctypes.c_int32), ("details", ctypes.c_int32),
                ("fastq_comments", ctypes.c_int32), ("only_mapped", ctypes.c_int32), ("use_ssw", ctypes.c_int32),
                ("rg_id", ctypes.c_char_p), ("rg", ctypes.POINTER(ctypes.c_char_p)), ("n_rg", ctypes.c_int32),
                ("max_secondary", ctypes.c_uint32), ("dropoff", ctypes.c_float),
                ("rescue_distance", ctypes.c_uint32), ("max_tries", ctypes.c_uint32), ("mcs", ctypes.c_uint32),
                ("chunk_size", ctypes.c_uint64), ("threads", ctypes.c_int32), ("batch_bases", ctypes.c_uint64),
                ("slots", ctypes.c_int32), ("match_score", ctypes.c_int32), ("mismatch_penalty", ctypes.c_int32),
                ("gap_open_penalty", ctypes.c_int32), ("gap_extend_penalty", ctypes.c_int32),
                ("end_bonus", ctypes.c_int32), ("xdrop", ctypes.c_int32), ("lookup_lanes", ctypes.c_int32),
                ("fused", ctypes.c_int32), ("seed_lanes", ctypes.c_int32)]


class Scores(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("match_score", "mismatch_penalty", "gap_open_penalty",
                                              "gap_extend_penalty", "end_bonus", "xdrop")]


class ExtendJob(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("padding", ctypes.c_uint32), ("anchor_offset", ctypes.c_uint64),
                ("n_anchors", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class NamExtension(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("found", "gapped", "score", "ref_start", "length", "query_start",
                                              "query_end", "edit_distance")] + \
               [("cigar_offset", ctypes.c_uint64), ("n_cigar", ctypes.c_uint32), ("extended", ctypes.c_uint32)]


class HitRec(ctypes.Structure):
    _fields_ = [("query_start", ctypes.c_uint64), ("query_end", ctypes.c_uint64), ("position", ctypes.c_uint64),
                ("hash", ctypes.c_uint64), ("hash_revcomp", ctypes.c_uint64), ("is_partial", ctypes.c_uint32),
                ("is_filtered", ctypes.c_uint32)]


class HitsDetails(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("full_not_found", "full_filtered", "full_found", "partial_not_found",
                                               "partial_filtered", "partial_found", "rescued")]


class QRand(ctypes.Structure):
    _fields_ = [("hash", ctypes.c_uint64), ("hash_revcomp", ctypes.c_uint64), ("start", ctypes.c_uint32),
                ("end", ctypes.c_uint32)]


class Nam(ctypes.Structure):
    _fields_ = [("nam_id", ctypes.c_uint32), ("ref_id", ctypes.c_uint32), ("ref_start", ctypes.c_uint32),
                ("ref_end", ctypes.c_uint32), ("query_start", ctypes.c_uint32), ("query_end", ctypes.c_uint32),
                ("matching_bases", ctypes.c_uint32), ("n_anchors", ctypes.c_uint32), ("score", ctypes.c_float),
                ("flags", ctypes.c_uint32), ("anchor_offset", ctypes.c_uint64)]


class Alignment(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("found", "score", "query_start", "query_end", "ref_start", "ref_end", "edit_distance", "n_cigar")]


EXPORTS = ["sa_last_error", "sa_version", "sa_ctx_create", "sa_ctx_destroy", "sa_seeding_params_for_read_length",
           "sa_index_build", "sa_index_build_device", "sa_index_save_sti", "sa_index_load_sti", "sa_index_download", "sa_index_upload", "sa_index_destroy", "sa_index_info", "sa_map_options_default",
           "sa_map_files", "sa_map_records", "sa_randstrobes_query_batch", "sa_find_nams_batch",
           "sa_find_nams_batch_noalloc", "sa_align_ssw_batch", "sa_extend_batch", "sa_find_hits_batch", "sa_chain_batch", "sa_map_batch", "sa_index_stats", "sa_index_kernel_times", "sa_reads_load", "sa_reads_load_device", "sa_reads_count", "sa_reads_record",
           "sa_reads_destroy", "sa_reads_upload", "sa_map_reads", "sa_hbm_copy_probe", "sa_server_create",
           "sa_server_destroy", "sa_server_submit", "sa_server_submit_reads", "sa_server_wait",
           "sa_server_get_stats", "sa_server_kernel_times", "sa_autotune", "sa_free"]


class RunStats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("upload_ms", ctypes.c_double), ("nam_ms", ctypes.c_double),
                ("extend_ms", ctypes.c_double), ("host_ms", ctypes.c_double), ("records", ctypes.c_uint64),
                ("requests", ctypes.c_uint64), ("rounds", ctypes.c_uint64), ("nams", ctypes.c_uint64),
                ("fused_batches", ctypes.c_uint64), ("spec_requests", ctypes.c_uint64), ("spec_used", ctypes.c_uint64),
                ("graph_replays", ctypes.c_uint64), ("fused_overflows", ctypes.c_uint64),
                ("replay_committed", ctypes.c_uint64), ("replay_probed", ctypes.c_uint64),
                ("replay_commit_ms", ctypes.c_double), ("replay_probe_ms", ctypes.c_double),
                ("replay_write_ms", ctypes.c_double), ("replay_accepted", ctypes.c_uint64)]


class Reads:
    """A read set parsed once (sa_reads_load), or parsed on the device of an
    index and left staged there (device=index: sa_reads_load_device; ingest_ms
    is then the device time of the parse)."""

    def __init__(self, reads1, reads2=None, device=None):
        load()
        self._p = ctypes.c_void_p()
        self.ingest_ms = None
        if device is not None:
            ms = ctypes.c_double()
            _check(_lib.sa_reads_load_device(device._p, reads1.encode(), reads2.encode() if reads2 else None,
                                             ctypes.byref(self._p), ctypes.byref(ms)))
            self.ingest_ms = ms.value
        else:
            _check(_lib.sa_reads_load(reads1.encode(), reads2.encode() if reads2 else None, ctypes.byref(self._p)))
        self.paired = reads2 is not None

    def __len__(self):
        return _lib.sa_reads_count(self._p)

    def record(self, i, mate=0):
        """(name, comment or None, sequence, qualities or None, end) of record i."""
        f = [ctypes.c_char_p() for _ in range(4)]
        end = ctypes.c_int()
        _check(_lib.sa_reads_record(self._p, ctypes.c_uint64(i), mate, *[ctypes.byref(x) for x in f],
                                    ctypes.byref(end)))
        return tuple(x.value for x in f) + (end.value,)

    def upload(self, index):
        """Stages the reads in the index's device memory (sa_reads_upload)."""
        _check(_lib.sa_reads_upload(index._p, self._p))

    def close(self):
        if self._p:
            _lib.sa_reads_destroy(self._p)
            self._p = None


def options(mode=0, eqx=False, rg_id=None, rg=(), threads=4, chunk_size=1000000, use_ssw=False, max_secondary=0,
            mcs=0, details=False, fastq_comments=False, only_mapped=False, batch_bases=0, slots=0, scores=None,
            xdrop=None, lookup_lanes=0, fused=False, seed_lanes=0):
    """Mapping options; scores = (match, mismatch, gap_open, gap_extend, end_bonus)
    as -A/-B/-O/-E/-L (defaults 2, 8, 12, 1, 10); lookup_lanes: lanes per read
    orientation in the lookup stage (0: default); fused: the fused batch pass
    (one replayed graph per batch) instead of the plain per-batch stages;
    seed_lanes: lanes per read in seeding, 1 or 64 (0: default)."""
    o = MapOptions()
    _lib.sa_map_options_default(ctypes.byref(o))
    if scores is not None:
        o.match_score, o.mismatch_penalty, o.gap_open_penalty, o.gap_extend_penalty, o.end_bonus = scores
    if xdrop is not None:
        o.xdrop = xdrop
    o.mode, o.eqx, o.details, o.fastq_comments, o.only_mapped = mode, int(eqx), int(details), int(fastq_comments), int(only_mapped)
    o.use_ssw, o.max_secondary, o.mcs, o.threads, o.chunk_size = int(use_ssw), max_secondary, mcs, threads, chunk_size
    o.batch_bases, o.slots, o.lookup_lanes, o.fused = batch_bases, slots, lookup_lanes, int(fused)
    o.seed_lanes = seed_lanes
    if rg_id is not None:
        o.rg_id = rg_id.encode()
    o._rg_keep = (ctypes.c_char_p * max(1, len(rg)))(*[x.encode() for x in rg])
    o.rg, o.n_rg = o._rg_keep, len(rg)
    return o


def load(path=LIB_PATH):
    """Loads the HIP library (raises SaError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SaError("libsa_hip.so not built (run make -C strobealign_amd/csrc)")
    lib = ctypes.CDLL(path)
    lib.sa_last_error.restype = ctypes.c_size_t
    lib.sa_version.restype = ctypes.c_char_p
    lib.sa_map_options_default.restype = None
    lib.sa_index_destroy.restype = None
    lib.sa_ctx_destroy.restype = None
    lib.sa_free.restype = None
    lib.sa_index_kernel_times.restype = ctypes.c_size_t
    lib.sa_reads_count.restype = ctypes.c_uint64
    lib.sa_reads_destroy.restype = None
    lib.sa_server_destroy.restype = None
    lib.sa_server_kernel_times.restype = ctypes.c_size_t
    _lib = lib
    return lib


def last_error():
    """Text of the library's last error on this thread."""
    buf = ctypes.create_string_buffer(4096)
    _lib.sa_last_error(buf, ctypes.c_size_t(4096))
    return buf.value.decode(errors="replace")


def _check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(4096)
        _lib.sa_last_error(buf, ctypes.c_size_t(4096))
        raise SaError("sa_hip error %d: %s" % (rc, buf.value.decode(errors="replace")))


class Context:
    """A device context (one GPU)."""

    def __init__(self, device=0):
        lib = load()
        self._p = ctypes.c_void_p()
        _check(lib.sa_ctx_create(device, ctypes.byref(self._p)))

    def close(self):
        if self._p:
            _lib.sa_ctx_destroy(self._p)
            self._p = None

    def align_ssw(self, pairs):
        """Aligner::align on the device for [(query, ref)] byte strings; returns
        [(found, score, qs, qe, rs, re, edit_distance, cigar_ops)]."""
        q = b"".join(p[0] for p in pairs)
        r = b"".join(p[1] for p in pairs)
        qo = [0]
        ro = [0]
        for a, b in pairs:
            qo.append(qo[-1] + len(a))
            ro.append(ro[-1] + len(b))
        n = len(pairs)
        out = (Alignment * max(1, n))()
        cig = ctypes.POINTER(ctypes.c_uint32)()
        coff = (ctypes.c_uint64 * (n + 1))()
        _check(_lib.sa_align_ssw_batch(self._p, q, (ctypes.c_uint64 * (n + 1))(*qo), r,
                                       (ctypes.c_uint64 * (n + 1))(*ro), ctypes.c_uint32(n), out,
                                       ctypes.byref(cig), coff))
        res = []
        for i in range(n):
            a = out[i]
            ops = [cig[x] for x in range(coff[i], coff[i + 1])]
            res.append((a.found, a.score, a.query_start, a.query_end, a.ref_start, a.ref_end, a.edit_distance, ops))
        _lib.sa_free(cig)
        return res

    # sa_extend_batch job kinds (include/sa_hip.h)
    EXT_SSW, EXT_HAMMING, EXT_PIECEWISE, EXT_GLOBAL, EXT_XDROP, EXT_XDROP_REV = range(6)

    def extend(self, jobs, scores=(2, 8, 12, 1, 10), xdrop=500, k=20):
        """sa_extend_batch for [(kind, query, ref, anchors, padding)] (anchors: a list
        of (ref_start, query_start) pairs, last anchor first, window coordinates);
        returns [(found, score, qs, qe, rs, re, edit_distance, cigar_ops)] with the
        score as a signed int."""
        n = len(jobs)
        q = b"".join(j[1] for j in jobs)
        r = b"".join(j[2] for j in jobs)
        qo, ro, flat = [0], [0], []
        js = (ExtendJob * max(1, n))()
        for i, (kind, a, b, anchors, padding) in enumerate(jobs):
            qo.append(qo[-1] + len(a))
            ro.append(ro[-1] + len(b))
            js[i].kind, js[i].padding = kind, padding
            js[i].anchor_offset, js[i].n_anchors = len(flat) // 2, len(anchors or ())
            for rs, qs in anchors or ():
                flat += [rs, qs]
        sc = Scores(*scores, xdrop)
        an = (ctypes.c_uint32 * max(1, len(flat)))(*flat)
        out = (Alignment * max(1, n))()
        cig = ctypes.POINTER(ctypes.c_uint32)()
        coff = (ctypes.c_uint64 * (n + 1))()
        _check(_lib.sa_extend_batch(self._p, ctypes.byref(sc), ctypes.c_uint32(k), js, ctypes.c_uint32(n), q,
                                    (ctypes.c_uint64 * (n + 1))(*qo), r, (ctypes.c_uint64 * (n + 1))(*ro), an,
                                    ctypes.c_uint64(len(flat) // 2), out, ctypes.byref(cig), coff))
        res = []
        for i in range(n):
            a = out[i]
            ops = [cig[x] for x in range(coff[i], coff[i + 1])]
            score = a.score - (1 << 32) if a.score >= (1 << 31) else a.score
            res.append((a.found, score, a.query_start, a.query_end, a.ref_start, a.ref_end, a.edit_distance, ops))
        _lib.sa_free(cig)
        return res

    def hbm_copy_gbps(self, nbytes=1 << 30, iters=20):
        """Read+write GB/s of a plain device copy (sa_hbm_copy_probe)."""
        v = ctypes.c_double()
        _check(_lib.sa_hbm_copy_probe(self._p, ctypes.c_uint64(nbytes), iters, ctypes.byref(v)))
        return v.value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Index:
    """Device-resident reference + randstrobe index (make_index, src/indexer.rs:18-170)."""

    def __init__(self, ctx, fasta, read_length, bits=0, threads=4, device_build=False, sti=None, upload=None):
        """device_build: run make_index on the device (sa_index_build_device);
        sti: load this .sti file instead of building (sa_index_load_sti);
        upload: (params, entries bytes, bucket starts, packed words, contig word
        offsets, contig lengths, contig names) for sa_index_upload."""
        self._p = ctypes.c_void_p()
        if upload is not None:
            prm, ent, bs, words, woff, lens, names = upload
            nm = (ctypes.c_char_p * max(1, len(names)))(*[x.encode() for x in names])
            _check(_lib.sa_index_upload(ctx._p, ctypes.byref(prm), ent, ctypes.c_uint64(len(ent) // 16),
                                        (ctypes.c_uint64 * len(bs))(*bs), (ctypes.c_uint64 * max(1, len(words)))(*words),
                                        ctypes.c_uint64(len(words)), (ctypes.c_uint64 * len(woff))(*woff),
                                        (ctypes.c_uint64 * max(1, len(lens)))(*lens), nm, ctypes.c_uint32(len(lens)),
                                        ctypes.byref(self._p)))
        elif sti is not None:
            _check(_lib.sa_index_load_sti(ctx._p, fasta.encode(), sti.encode(), ctypes.c_uint64(read_length), bits,
                                          ctypes.byref(self._p)))
        elif device_build:
            _check(_lib.sa_index_build_device(ctx._p, fasta.encode(), ctypes.c_uint64(read_length), bits,
                                              ctypes.byref(self._p)))
        else:
            _check(_lib.sa_index_build(ctx._p, fasta.encode(), ctypes.c_uint64(read_length), bits, threads,
                                       ctypes.byref(self._p)))

    def close(self):
        if self._p:
            _lib.sa_index_destroy(self._p)
            self._p = None

    def params(self):
        p = SeedingParams()
        n = ctypes.c_uint64()
        c = ctypes.c_uint32()
        _check(_lib.sa_index_info(self._p, ctypes.byref(p), ctypes.byref(n), ctypes.byref(c)))
        return p, n.value, c.value

    def save_sti(self, path, read_length):
        """Writes the index as a .sti file (sa_index_save_sti)."""
        _check(_lib.sa_index_save_sti(self._p, ctypes.c_uint64(read_length), path.encode()))

    def download(self):
        """(entries as bytes, 16 per entry; bucket starts as a list) copied from the device."""
        p, n, _ = self.params()
        ent = ctypes.create_string_buffer(max(1, 16 * n))
        bs = (ctypes.c_uint64 * ((1 << p.bits) + 1))()
        _check(_lib.sa_index_download(self._p, ent, bs))
        return ent.raw[:16 * n], list(bs)

    def map_files(self, reads1, reads2=None, mode=0, eqx=False, rg_id=None, rg=(), threads=4, chunk_size=1000000,
                  use_ssw=False, max_secondary=0, mcs=0, details=False, fastq_comments=False, only_mapped=False,
                  batch_bases=0, scores=None, xdrop=None, lookup_lanes=0, fused=False, slots=0, seed_lanes=0):
        o = options(mode=mode, eqx=eqx, rg_id=rg_id, rg=rg, threads=threads, chunk_size=chunk_size, use_ssw=use_ssw,
                    max_secondary=max_secondary, mcs=mcs, details=details, fastq_comments=fastq_comments,
                    only_mapped=only_mapped, batch_bases=batch_bases, scores=scores, xdrop=xdrop,
                    lookup_lanes=lookup_lanes, fused=fused, slots=slots, seed_lanes=seed_lanes)
        out = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _check(_lib.sa_map_files(self._p, ctypes.byref(o), reads1.encode(),
                                 reads2.encode() if reads2 else None, ctypes.byref(out), ctypes.byref(n)))
        s = ctypes.string_at(out, n.value)
        _lib.sa_free(out)
        return s

    def autotune(self, reads, opts, sample_pairs=0):
        """Chooses opts.lookup_lanes / opts.slots / opts.batch_bases by timing the
        read set's leading sample_pairs records (sa_autotune); returns the trials."""
        import json
        buf = ctypes.create_string_buffer(8192)
        _check(_lib.sa_autotune(self._p, reads._p, ctypes.c_uint64(sample_pairs), ctypes.byref(opts), buf,
                                ctypes.c_size_t(8192)))
        return json.loads(buf.value.decode())

    def map_reads(self, reads, opts, keep_output=True):
        """Maps a Reads set; returns (output bytes or None, RunStats)."""
        st = RunStats()
        if keep_output:
            out = ctypes.c_void_p()
            n = ctypes.c_size_t()
            _check(_lib.sa_map_reads(self._p, ctypes.byref(opts), reads._p, ctypes.byref(out), ctypes.byref(n),
                                     ctypes.byref(st)))
            s = ctypes.string_at(out, n.value)
            _lib.sa_free(out)
            return s, st
        _check(_lib.sa_map_reads(self._p, ctypes.byref(opts), reads._p, None, None, ctypes.byref(st)))
        return None, st

    @staticmethod
    def _pack(seqs):
        data = b"".join(seqs)
        off = [0]
        for s in seqs:
            off.append(off[-1] + len(s))
        return data, (ctypes.c_uint64 * len(off))(*off)

    def randstrobes(self, seqs):
        data, off = self._pack(seqs)
        n = len(seqs)
        tot = max(1, off[n])
        fwd, rc = (QRand * tot)(), (QRand * tot)()
        counts = (ctypes.c_uint32 * (2 * n + 1))()
        _check(_lib.sa_randstrobes_query_batch(self._p, data, off, ctypes.c_uint32(n), fwd, rc, counts))
        res = []
        for i in range(n):
            o = off[i]
            res.append(([(fwd[o + x].hash, fwd[o + x].hash_revcomp, fwd[o + x].start, fwd[o + x].end)
                         for x in range(counts[2 * i])],
                        [(rc[o + x].hash, rc[o + x].hash_revcomp, rc[o + x].start, rc[o + x].end)
                         for x in range(counts[2 * i + 1])]))
        return res

    def map_batch(self, seqs, opts, speculative=3, stream=None):
        """sa_map_batch: per read, its NAMs (as nams() gives them) and for each NAM
        None (not extended) or (found, gapped, score, ref_start, length,
        query_start, query_end, edit_distance, cigar_ops)."""
        data, off = self._pack(seqs)
        n = len(seqs)
        p = ctypes.POINTER(Nam)()
        e = ctypes.POINTER(NamExtension)()
        c = ctypes.POINTER(ctypes.c_uint32)()
        noff = (ctypes.c_uint64 * (n + 1))()
        _check(_lib.sa_map_batch(self._p, ctypes.byref(opts), data, off, ctypes.c_uint32(n),
                                 ctypes.c_uint32(speculative), ctypes.c_void_p(stream), ctypes.byref(p), noff,
                                 ctypes.byref(e), ctypes.byref(c)))
        nams, exts = [], []
        for i in range(n):
            nams.append([(p[x].nam_id, p[x].ref_id, p[x].ref_start, p[x].ref_end, p[x].query_start, p[x].query_end,
                          p[x].matching_bases, p[x].n_anchors, p[x].score, p[x].flags & 1)
                         for x in range(noff[i], noff[i + 1])])
            row = []
            for x in range(noff[i], noff[i + 1]):
                a = e[x]
                if not a.extended:
                    row.append(None)
                    continue
                ops = [c[y] for y in range(a.cigar_offset, a.cigar_offset + a.n_cigar)] if a.found else []
                row.append((a.found, a.gapped, a.score, a.ref_start, a.length, a.query_start, a.query_end,
                            a.edit_distance, ops))
            exts.append(row)
        for q in (p, e, c):
            _lib.sa_free(q)
        return nams, exts

    def find_hits(self, seqs, mcs=0, rescue_distance=100):
        """sa_find_hits_batch: per read orientation (2 per read, forward first) a
        list of hits (query_start, query_end, position, hash, hash_revcomp,
        is_partial, is_filtered) and the details tuple."""
        data, off = self._pack(seqs)
        n = len(seqs)
        p = ctypes.POINTER(HitRec)()
        hoff = (ctypes.c_uint64 * (2 * n + 1))()
        det = (HitsDetails * max(1, 2 * n))()
        _check(_lib.sa_find_hits_batch(self._p, data, off, ctypes.c_uint32(n), mcs, rescue_distance, ctypes.byref(p),
                                       hoff, det))
        hits = [[(p[x].query_start, p[x].query_end, p[x].position, p[x].hash, p[x].hash_revcomp, p[x].is_partial,
                  p[x].is_filtered) for x in range(hoff[t], hoff[t + 1])] for t in range(2 * n)]
        dets = [tuple(getattr(det[t], f) for f, _ in HitsDetails._fields_) for t in range(2 * n)]
        _lib.sa_free(p)
        return hits, dets

    def chain_raw(self, seqs, hits, dets):
        """sa_chain_batch over hits and details as find_hits returns them: (NAM
        records as bytes, the n + 1 offsets), like nams_raw."""
        data, off = self._pack(seqs)
        n = len(seqs)
        flat = [h for t in hits for h in t]
        arr = (HitRec * max(1, len(flat)))(*[HitRec(*h) for h in flat])
        hoff = [0]
        for t in hits:
            hoff.append(hoff[-1] + len(t))
        det = (HitsDetails * max(1, 2 * n))(*[HitsDetails(*d) for d in dets])
        p = ctypes.POINTER(Nam)()
        noff = (ctypes.c_uint64 * (n + 1))()
        _check(_lib.sa_chain_batch(self._p, data, off, ctypes.c_uint32(n), arr, (ctypes.c_uint64 * (2 * n + 1))(*hoff),
                                   det, ctypes.byref(p), noff))
        raw = ctypes.string_at(p, 48 * noff[n])
        _lib.sa_free(p)
        return raw, list(noff)

    def nams_raw(self, seqs, mcs=0, rescue_distance=100):
        """sa_find_nams_batch: (the NAM records as bytes, 48 per record; the n + 1 offsets)."""
        data, off = self._pack(seqs)
        n = len(seqs)
        p = ctypes.POINTER(Nam)()
        noff = (ctypes.c_uint64 * (n + 1))()
        _check(_lib.sa_find_nams_batch(self._p, data, off, ctypes.c_uint32(n), mcs, rescue_distance,
                                       ctypes.byref(p), noff))
        raw = ctypes.string_at(p, 48 * noff[n])
        _lib.sa_free(p)
        return raw, list(noff)

    def nams(self, seqs, mcs=0, rescue_distance=100):
        data, off = self._pack(seqs)
        n = len(seqs)
        p = ctypes.POINTER(Nam)()
        noff = (ctypes.c_uint64 * (n + 1))()
        _check(_lib.sa_find_nams_batch(self._p, data, off, ctypes.c_uint32(n), mcs, rescue_distance,
                                       ctypes.byref(p), noff))
        res = []
        for i in range(n):
            res.append([(p[x].nam_id, p[x].ref_id, p[x].ref_start, p[x].ref_end, p[x].query_start, p[x].query_end,
                         p[x].matching_bases, p[x].n_anchors, p[x].score, p[x].flags & 1)
                        for x in range(noff[i], noff[i + 1])])
        _lib.sa_free(p)
        return res

    def kernel_times(self):
        """{kernel: (total_ms, launches, items, algorithmic_bytes, dp_cells)} timed with HIP events."""
        import json
        n = _lib.sa_index_kernel_times(self._p, None, ctypes.c_size_t(0))
        buf = ctypes.create_string_buffer(n + 1)
        _lib.sa_index_kernel_times(self._p, buf, ctypes.c_size_t(n + 1))
        return {k: tuple(v) for k, v in json.loads(buf.value.decode()).items()}

    def kernel_ms(self):
        v = ctypes.c_double()
        _check(_lib.sa_index_stats(self._p, ctypes.byref(v)))
        return v.value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "batches", "batch_requests",
                                                "max_requests_per_batch", "arena_bytes", "arena_peak")] + \
               [("busy_ms", ctypes.c_double)]


class Server:
    """Serving loop with continuous batching (sa_server_*): requests are read sets
    submitted as they arrive; each one's output equals mapping it alone."""

    def __init__(self, index, opts, arena_bytes=0):
        self._p = ctypes.c_void_p()
        self._index = index
        _check(_lib.sa_server_create(index._p, ctypes.byref(opts), ctypes.c_uint64(arena_bytes), ctypes.byref(self._p)))

    def submit(self, reads, first, count):
        """Records [first, first + count) of a Reads set as one request; returns its ticket."""
        t = ctypes.c_uint64()
        _check(_lib.sa_server_submit_reads(self._p, reads._p, ctypes.c_uint64(first), ctypes.c_uint64(count),
                                           ctypes.byref(t)))
        return t.value

    def submit_records(self, recs1, recs2=None):
        """recs: lists of (name, seq, qual) bytes (qual may be None); returns the ticket."""
        n = len(recs1)

        def arrays(recs):
            names = (ctypes.c_char_p * max(1, n))(*[r[0] for r in recs])
            seqs = (ctypes.c_char_p * max(1, n))(*[r[1] for r in recs])
            quals = None
            if n and recs[0][2] is not None:
                quals = (ctypes.c_char_p * n)(*[r[2] for r in recs])
            return names, seqs, quals
        a1 = arrays(recs1)
        a2 = arrays(recs2) if recs2 is not None else (None, None, None)
        t = ctypes.c_uint64()
        _check(_lib.sa_server_submit(self._p, ctypes.c_uint64(n), a1[0], a1[1], a1[2], a2[0], a2[1], a2[2],
                                     ctypes.byref(t)))
        return t.value

    def wait(self, ticket, keep_output=True):
        """Blocks until the request is done; returns (output bytes or None, latency ms)."""
        lat = ctypes.c_double()
        if keep_output:
            out = ctypes.c_void_p()
            n = ctypes.c_size_t()
            _check(_lib.sa_server_wait(self._p, ctypes.c_uint64(ticket), ctypes.byref(out), ctypes.byref(n),
                                       ctypes.byref(lat)))
            s = ctypes.string_at(out, n.value)
            _lib.sa_free(out)
            return s, lat.value
        _check(_lib.sa_server_wait(self._p, ctypes.c_uint64(ticket), None, None, ctypes.byref(lat)))
        return None, lat.value

    def stats(self):
        st = ServerStats()
        _check(_lib.sa_server_get_stats(self._p, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def kernel_times(self):
        import json
        n = _lib.sa_server_kernel_times(self._p, None, ctypes.c_size_t(0))
        buf = ctypes.create_string_buffer(n + 1)
        _lib.sa_server_kernel_times(self._p, buf, ctypes.c_size_t(n + 1))
        return {k: tuple(v) for k, v in json.loads(buf.value.decode()).items()}

    def close(self):
        if self._p:
            _lib.sa_server_destroy(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
