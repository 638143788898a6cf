// Residency and device-side whole-substate operations of the engine.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "engine.hpp"

namespace calhip {

int ensure_stream(ModelX* m)
{
	if (m->stream) return 0;
	int n = 0;
	if (CALHIP_CHECK(hipGetDeviceCount(&n))) return -1;
	if (n <= 0) {
		set_error(4, "no GPU is visible to the HIP runtime");
		return -1;
	}
	if (CALHIP_CHECK(hipGetDevice(&m->device))) return -1;
	// default priority: work on streams of other priorities is avoided
	return CALHIP_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
}

int ensure_device_buffers(ModelX* m, SubX* s)
{
	size_t bytes = m->cells() * (size_t)s->elem;
	if (!s->d_cur && CALHIP_CHECK(hipMalloc(&s->d_cur, bytes))) return -1;
	if (!s->single && !s->d_nxt && CALHIP_CHECK(hipMalloc(&s->d_nxt, bytes))) return -1;
	return 0;
}

static int ensure_active_buffers(ModelX* m)
{
	if (!m->active()) return 0;
	size_t n = m->cells();
	// rounded up to whole 16-byte words: kernels read flags 16 at a time and
	// set one with a 32-bit atomic on the word that holds it
	if (!m->d_flags && CALHIP_CHECK(hipMalloc(&m->d_flags, (n + 15) & ~(size_t)15))) return -1;
	if (!m->d_cells && CALHIP_CHECK(hipMalloc(&m->d_cells, n * sizeof(int)))) return -1;
	if (!m->d_count && CALHIP_CHECK(hipMalloc(&m->d_count, 4 * sizeof(int)))) return -1;
	return 0;
}

// HOST -> DEVICE: upload both layers of every substate and the active set.
int ensure_device(ModelX* m)
{
	if (m->phase == Phase::Device) return 0;
	if (ensure_stream(m)) return -1;
	CALModel2D* p = &m->pub;
	size_t n = m->cells();
	for (SubX* s : m->subs) {
		if (s->host_only) continue;
		if (ensure_device_buffers(m, s)) return -1;
		size_t bytes = n * (size_t)s->elem;
		s->tstate = Transient::None;
		if (CALHIP_CHECK(hipMemcpyAsync(s->d_cur, s->h_cur(), bytes, hipMemcpyHostToDevice, m->stream)))
			return -1;
		if (!s->single) {
			if (memcmp(s->h_cur(), s->h_nxt(), bytes) == 0) {
				s->nxt_stale = true;
				s->nxt_dirty = false;
			} else {
				if (CALHIP_CHECK(hipMemcpyAsync(s->d_nxt, s->h_nxt(), bytes, hipMemcpyHostToDevice,
				                                m->stream)))
					return -1;
				s->nxt_stale = false;
				s->nxt_dirty = true;
			}
		}
	}
	if (m->active()) {
		if (ensure_active_buffers(m)) return -1;
		if (CALHIP_CHECK(hipMemcpyAsync(m->d_flags, p->A.flags, n, hipMemcpyHostToDevice, m->stream)))
			return -1;
		// the device list mirrors A.cells (linear indices); count 0 <=> NULL list
		std::vector<int> lin;
		int cnt;
		if (m->pub3) {
			const CALModel3D* q = m->pub3;
			cnt = q->A.cells ? q->A.size_current : 0;
			lin.resize((size_t)cnt + 1);
			for (int k = 0; k < cnt; ++k)
				lin[k] = (q->A.cells[k].k * q->rows + q->A.cells[k].i) * q->columns + q->A.cells[k].j;
		} else {
			cnt = p->A.cells ? p->A.size_current : 0;
			lin.resize((size_t)cnt + 1);
			for (int k = 0; k < cnt; ++k) lin[k] = p->A.cells[k].i * p->columns + p->A.cells[k].j;
		}
		if (cnt && CALHIP_CHECK(hipMemcpyAsync(m->d_cells, lin.data(), sizeof(int) * (size_t)cnt,
		                                       hipMemcpyHostToDevice, m->stream)))
			return -1;
		int hc[4] = {cnt, 0, 0, 0};
		if (CALHIP_CHECK(hipMemcpyAsync(m->d_count, hc, sizeof(hc), hipMemcpyHostToDevice, m->stream)))
			return -1;
		if (CALHIP_CHECK(hipStreamSynchronize(m->stream))) return -1;
		int pending = 0;
		for (int t = 0; t < p->A.num_threads; ++t) pending += p->A.size_next[t];
		m->flags_dirty = pending != 0;
		for (int t = 0; t < p->A.num_threads; ++t) p->A.size_next[t] = 0;
	}
	if (CALHIP_CHECK(hipStreamSynchronize(m->stream))) return -1;
	m->phase = Phase::Device;
	m->stats.uploads++;
	// the host mirrors hold what was just uploaded
	device_changed(m);
	for (SubX* s : m->subs) s->host_epoch = m->dev_epoch;
	if (m->prog && m->prog->after_upload(m)) return -1;
	return 0;
}

// DEVICE -> HOST: materialise transient substates, download everything.
int ensure_host(ModelX* m)
{
	if (m->phase == Phase::Host) return 0;
	CALModel2D* p = &m->pub;
	size_t n = m->cells();
	if (m->prog && m->prog->sync_active(m)) return -1;
	for (SubX* s : m->subs) {
		if (s->host_only) continue;
		if (s->tstate != Transient::None && s->tstate != Transient::Zero && m->prog) {
			if (m->prog->materialise(m, s)) return -1;
			m->stats.materialisations++;
		}
	}
	for (SubX* s : m->subs) {
		if (s->host_only) continue;
		size_t bytes = n * (size_t)s->elem;
		if (s->tstate == Transient::Zero) {
			memset(s->h_cur(), 0, bytes);
			if (!s->single) memset(s->h_nxt(), 0, bytes);
			continue;
		}
		if (CALHIP_CHECK(hipMemcpyAsync(s->h_cur(), s->d_cur, bytes, hipMemcpyDeviceToHost, m->stream)))
			return -1;
		if (!s->single && !s->nxt_stale &&
		    CALHIP_CHECK(hipMemcpyAsync(s->h_nxt(), s->d_nxt, bytes, hipMemcpyDeviceToHost, m->stream)))
			return -1;
	}
	int cnt = 0;
	std::vector<int> lin;
	if (m->active()) {
		if (CALHIP_CHECK(hipMemcpyAsync(p->A.flags, m->d_flags, n, hipMemcpyDeviceToHost, m->stream)))
			return -1;
		if (CALHIP_CHECK(hipMemcpyAsync(&cnt, m->d_count, sizeof(int), hipMemcpyDeviceToHost, m->stream)))
			return -1;
	}
	if (CALHIP_CHECK(hipStreamSynchronize(m->stream))) return -1;
	for (SubX* s : m->subs)
		if (!s->host_only && !s->single && s->nxt_stale && s->tstate != Transient::Zero)
			memcpy(s->h_nxt(), s->h_cur(), n * (size_t)s->elem);
	if (m->active()) {
		lin.resize((size_t)cnt + 1);
		if (cnt && CALHIP_CHECK(hipMemcpy(lin.data(), m->d_cells, sizeof(int) * (size_t)cnt,
		                                  hipMemcpyDeviceToHost)))
			return -1;
		if (m->pub3) {
			// OpenCAL's 3D list is in (i, j, k) order (cal3D.c:412-424)
			CALModel3D* q = m->pub3;
			free(q->A.cells);
			q->A.cells = nullptr;
			q->A.size_current = cnt;
			if (cnt) {
				q->A.cells = (CALCell3D*)malloc(sizeof(CALCell3D) * (size_t)cnt);
				const long long plane = (long long)q->rows * q->columns;
				for (int k = 0; k < cnt; ++k) {
					const long long r = lin[k] % plane;
					q->A.cells[k].k = (int)(lin[k] / plane);
					q->A.cells[k].i = (int)(r / q->columns);
					q->A.cells[k].j = (int)(r % q->columns);
				}
				std::sort(q->A.cells, q->A.cells + cnt, [](const CALCell3D& a, const CALCell3D& b) {
					return a.i != b.i ? a.i < b.i : (a.j != b.j ? a.j < b.j : a.k < b.k);
				});
			}
		} else {
			free(p->A.cells);
			p->A.cells = nullptr;
			p->A.size_current = cnt;
			if (cnt) {
				p->A.cells = (CALCell2D*)malloc(sizeof(CALCell2D) * (size_t)cnt);
				for (int k = 0; k < cnt; ++k) {
					p->A.cells[k].i = lin[k] / p->columns;
					p->A.cells[k].j = lin[k] % p->columns;
				}
			}
		}
		// flags changed since the last rebuild show up as a pending delta
		if (m->flags_dirty) {
			int set = 0;
			for (size_t c = 0; c < n; ++c) set += p->A.flags[c] ? 1 : 0;
			p->A.size_next[0] = set - cnt;
		}
	}
	for (SubX* s : m->subs) {
		s->tstate = Transient::None;
		s->nxt_stale = false;
		s->nxt_dirty = false;
	}
	m->phase = Phase::Host;
	m->stats.downloads++;
	return 0;
}

int host_read(ModelX* m, void* q)
{
	SubX* s = S(q);
	if (m->phase == Phase::Host || s->host_only) return 0;
	if (s->tstate != Transient::None) return ensure_host(m);   // logical content is not in the buffer
	if (s->host_epoch == m->dev_epoch) return 0;
	if (m->prog && m->prog->before_read(m)) return -1;
	if (CALHIP_CHECK(hipMemcpyAsync(s->h_cur(), s->d_cur, m->cells() * (size_t)s->elem, hipMemcpyDeviceToHost,
	                                m->stream)) ||
	    CALHIP_CHECK(hipStreamSynchronize(m->stream)))
		return -1;
	s->host_epoch = m->dev_epoch;
	m->stats.downloads++;
	return 0;
}

void flip(SubX* s)
{
	void* t = s->d_cur;
	s->d_cur = s->d_nxt;
	s->d_nxt = t;
	s->nxt_dirty = false;
	s->nxt_stale = true;
}

int materialise_next(ModelX* m, SubX* s)
{
	if (s->single || !s->nxt_stale) return 0;
	if (k::copy(m, s->d_nxt, s->d_cur, m->cells() * (size_t)s->elem)) return -1;
	s->nxt_stale = false;
	return 0;
}

// calInitSubstate(Next)2D* on the device.
int device_fill(ModelX* m, SubX* s, double v, bool both)
{
	device_changed(m);
	if (s->tstate != Transient::None && m->prog) {
		int r = m->prog->transient_init(m, s, v, both);
		if (r != 0) return r < 0 ? -1 : 0;
		if (m->prog->materialise(m, s)) return -1;
		s->tstate = Transient::None;
		m->stats.materialisations++;
	}
	if (m->prog && m->prog->before_modify(m, s)) return -1;
	bool partial = m->active();   // over the active list when it is non-empty
	if (partial && m->prog && m->prog->sync_active(m)) return -1;
	if (s->single) return k::fill(m, s->d_cur, s->elem, v, partial);
	if (both) {
		if (k::fill(m, s->d_cur, s->elem, v, partial)) return -1;
		if (s->nxt_stale) return 0;  // logical next == current still holds
		return k::fill(m, s->d_nxt, s->elem, v, partial);
	}
	if (partial && materialise_next(m, s)) return -1;
	if (k::fill(m, s->d_nxt, s->elem, v, partial)) return -1;
	s->nxt_stale = false;
	s->nxt_dirty = true;
	return 0;
}

int device_rebuild_active(ModelX* m)
{
	if (!m->active() || !m->flags_dirty) return 0;
	device_changed(m);
	if (k::rebuild_active(m)) return -1;
	m->flags_dirty = false;
	m->stats.rebuilds++;
	return 0;
}

int Program::device_update_hook(ModelX* m) { return device_update(m); }

// calUpdateSubstate2D* of one written next layer on the device
// (cal2D.c:694-713): in active mode next -> current at the cells of the
// active list (the whole grid while it is empty), so that a cell written by
// a process and then removed from the set keeps its current value as in the
// reference; otherwise a flip of the two buffers.
int device_update_sub(ModelX* m, SubX* s)
{
	if (s->host_only || s->single || !s->nxt_dirty) return 0;
	if (!m->active()) {
		flip(s);
		return 0;
	}
	if (m->prog && m->prog->sync_active(m)) return -1;
	if (k::copy_active(m, s->d_cur, s->d_nxt, s->elem)) return -1;
	s->nxt_dirty = false;
	s->nxt_stale = false;   // the next layer keeps its own values outside the set
	return 0;
}

// calUpdate2D on the device: rebuild, then update every written next layer.
int device_update(ModelX* m)
{
	device_changed(m);
	if (device_rebuild_active(m)) return -1;
	for (SubX* s : m->subs)
		if (device_update_sub(m, s)) return -1;
	if (m->prog) m->prog->transient_update(m);
	return 0;
}

// ------------------------------------------------------------- timing

void timer_begin(ModelX* m, hipEvent_t* ev0)
{
	*ev0 = nullptr;
	if (!m->timer.enabled) return;
	hipEventCreate(ev0);
	hipEventRecord(*ev0, m->stream);
}

void timer_end(ModelX* m, hipEvent_t ev0)
{
	if (!m->timer.enabled || !ev0) return;
	hipEvent_t ev1;
	hipEventCreate(&ev1);
	hipEventRecord(ev1, m->stream);
	m->timer.pending.push_back({ev0, ev1});
	if (m->timer.pending.size() > 512) timer_collect(m);
}

void timer_collect(ModelX* m)
{
	for (auto& p : m->timer.pending) {
		float ms = 0.f;
		hipEventSynchronize(p.second);
		hipEventElapsedTime(&ms, p.first, p.second);
		m->timer.total_ms += ms;
		m->timer.launches++;
		hipEventDestroy(p.first);
		hipEventDestroy(p.second);
	}
	m->timer.pending.clear();
}

}  // namespace calhip

using namespace calhip;

extern "C" int calhipSyncHost2D(struct CALModel2D* ca2D) { return ensure_host(X(ca2D)); }

extern "C" int calhipSynchronize2D(struct CALModel2D* ca2D)
{
	ModelX* m = X(ca2D);
	if (m->prog && m->prog->flush(m)) return -1;
	if (!m->stream) return 0;
	return CALHIP_CHECK(hipStreamSynchronize(m->stream));
}

extern "C" int calhipIsOnDevice2D(struct CALModel2D* ca2D) { return X(ca2D)->phase == Phase::Device; }

extern "C" int calhipGetStats2D(struct CALModel2D* ca2D, struct calhipStats* out)
{
	ModelX* m = X(ca2D);
	if (m->prog && m->prog->flush(m)) return -1;
	*out = m->stats;
	return 0;
}

extern "C" int calhipEnableKernelTiming2D(struct CALModel2D* ca2D, int enable)
{
	ModelX* m = X(ca2D);
	if (m->prog && m->prog->flush(m)) return -1;
	if (!enable) timer_collect(m);
	if (m->timer.region_start) {
		hipEventDestroy(m->timer.region_start);
		m->timer.region_start = nullptr;
	}
	m->timer.region = enable == 2;
	m->timer.enabled = enable == 1;
	if (enable) {
		m->timer.total_ms = 0.0;
		m->timer.launches = 0;
	}
	if (m->timer.region) {
		if (ensure_stream(m) || CALHIP_CHECK(hipEventCreate(&m->timer.region_start)) ||
		    CALHIP_CHECK(hipEventRecord(m->timer.region_start, m->stream)))
			return -1;
		m->timer.region_launches0 = m->stats.launches;
	}
	return 0;
}

extern "C" int calhipGetKernelTime2D(struct CALModel2D* ca2D, double* total_ms, long long* launches)
{
	ModelX* m = X(ca2D);
	if (m->prog && m->prog->flush(m)) return -1;
	if (m->timer.region && m->timer.region_start) {
		hipEvent_t end;
		float ms = 0.f;
		if (CALHIP_CHECK(hipEventCreate(&end))) return -1;
		const int rc = CALHIP_CHECK(hipEventRecord(end, m->stream)) || CALHIP_CHECK(hipEventSynchronize(end)) ||
		               CALHIP_CHECK(hipEventElapsedTime(&ms, m->timer.region_start, end));
		hipEventDestroy(end);
		if (rc) return -1;
		*total_ms = ms;
		*launches = m->stats.launches - m->timer.region_launches0;
		return 0;
	}
	timer_collect(m);
	*total_ms = m->timer.total_ms;
	*launches = m->timer.launches;
	return 0;
}
