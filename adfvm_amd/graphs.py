"""Primal + adjoint sweep with HIP-graph segments.

For fixed-dt runs the work between two checkpoints is the same every time:
`interval` primal steps forward (objective accumulated on the device), and
for the reverse the same steps re-run keeping their start states and tapes,
then `interval` adjoint steps.  Each of the two is captured once into a HIP
graph (torch.cuda.CUDAGraph) and replayed per segment, so a segment costs
one graph launch instead of a few hundred kernel launches -- what matters on
meshes small enough for launch overhead to show (the cylinder O-grid, small
ensemble members).  The reverse sweep is the uniform checkpoint schedule of
apps/adjoint.py:205-217 (checkpoints every writeInterval steps).

With whole=True the entire sweep for a fixed number of steps -- the forward
run writing its checkpoints into a preallocated buffer, then every reverse
segment -- is one captured graph, replayed as a single launch per sweep.

With overlap (the default when a second set of segment buffers fits in
device memory), the re-run of segment s-1 goes on a second stream while the
adjoint steps of segment s run: the two use separate stage-state and tape
buffers and disjoint context work buffers (the primal's face-flux records,
the adjoint's face and cell adjoint records), and events order each re-run
after the adjoint steps that last read its buffers and before the adjoint
steps that read them next.  Every kernel computes what it computed in
sequence: the results are bit-identical.
"""
import numpy as np


class GraphSweep(object):
    """Buffers and captured graphs for repeated sweeps of one context at one
    dt, checkpoint interval and objective weight.  The graphs are captured
    on the first run (after one eager warm-up segment on a side stream) and
    replayed by every later run.  With bc_grad, the boundary-value
    gradient d(obja * J)/d(BC values) is accumulated into self.bc_grad
    ((5, nBoundaryFaces), Context.bc_param locates a parameter in it)."""

    def __init__(self, ctx, dt, interval, obja=1.0, graphs=True, tapes=True, source=None, bc_grad=False,
                 whole=False, overlap=None):
        torch = ctx.torch
        self.torch, self.ctx = torch, ctx
        self.dt, self.interval, self.obja, self.source = float(dt), int(interval), float(obja), source
        self.graphs = graphs
        self.whole = whole                  # the whole sweep as one graph (needs graphs)
        self._whole = None                  # (n_steps, graph, checkpoint buffer, input buffer)
        nseg = getattr(ctx, 'nSeg', 1)
        shape = (5, ctx.nIC)
        self.u = torch.empty(shape, dtype=ctx.tdtype, device='cuda')
        self.obj = torch.zeros((interval, nseg), dtype=torch.float64, device='cuda')
        self.states = torch.empty((interval,) + shape, dtype=ctx.tdtype, device='cuda')
        self.tps = ctx.new_tapes(interval) if tapes else None
        self.ua = torch.zeros(shape, dtype=ctx.tdtype, device='cuda')
        self.J = torch.zeros(nseg, dtype=torch.float64, device='cuda')
        self.bc_grad = ctx.bc_state() if bc_grad else None
        self._fwd = self._rev = None
        # overlapped reverse: a second set of segment buffers, the re-run's
        # working state and its stream
        if overlap is None:
            need = self.states.numel() * self.states.element_size() + \
                (self.tps.numel() * self.tps.element_size() if tapes else 0)
            overlap = tapes and need < torch.cuda.mem_get_info()[0] - (16 << 30)
        if overlap and not tapes:
            raise ValueError('overlap needs tapes (the adjoint steps read the re-run\'s tapes)')
        self.overlap = bool(overlap)
        self._bufs = [(self.states, self.tps)]
        if self.overlap:
            self._bufs.append((torch.empty_like(self.states), ctx.new_tapes(interval)))
            self.u_side = torch.empty(shape, dtype=ctx.tdtype, device='cuda')
            self._side = torch.cuda.Stream()
        self._rerun_g = self._adj_g = None
        self._captured_for = None        # the context's objective state the graphs were captured with
        self.keep_final = False          # set to keep the state after the forward sweep (u_final)
        self.u_final = None

    def _forward_segment(self):
        ctx, u, dt = self.ctx, self.u, self.dt
        for k in range(self.interval):
            ctx.primal_step(u, dt, source=self.source, obj_out=self.obj[k])
            self.J.add_(self.obj[k])            # step order, as a single run accumulates

    def _reverse_segment(self):
        ctx, u, dt, tps = self.ctx, self.u, self.dt, self.tps
        for k in range(self.interval):
            self.states[k].copy_(u)
            ctx.primal_step(u, dt, source=self.source, tape=None if tps is None else tps[k])
        for k in reversed(range(self.interval)):
            ctx.adjoint_step(self.states[k], dt, self.ua, obja=self.obja, source=self.source,
                             bc_grad=self.bc_grad, tape=None if tps is None else tps[k])

    def _rerun_segment(self, b, u):
        """The segment's steps from the state in u (advanced in place),
        keeping start states and tapes in buffer set b."""
        ctx, dt = self.ctx, self.dt
        states, tps = self._bufs[b]
        for k in range(self.interval):
            states[k].copy_(u)
            ctx.primal_step(u, dt, source=self.source, tape=tps[k])

    def _adjoint_segment(self, b):
        ctx, dt = self.ctx, self.dt
        states, tps = self._bufs[b]
        for k in reversed(range(self.interval)):
            ctx.adjoint_step(states[k], dt, self.ua, obja=self.obja, source=self.source,
                             bc_grad=self.bc_grad, tape=tps[k])

    def _reverse_overlapped(self, checkpoints, run_rerun, run_adj):
        """The reverse segments, last first, each segment's adjoint steps
        alongside the previous segment's re-run (buffer sets alternate)."""
        torch = self.torch
        main, side = torch.cuda.current_stream(), self._side
        nseg = len(checkpoints)
        self.u_side.copy_(checkpoints[nseg - 1])
        run_rerun((nseg - 1) % 2)
        for s in range(nseg - 1, -1, -1):
            if s > 0:
                # after everything issued so far: the re-run of segment s
                # (joined below in the previous round) and the adjoint steps
                # of s+1, the last readers of buffer set (s-1) % 2
                side.wait_stream(main)
                with torch.cuda.stream(side):
                    self.u_side.copy_(checkpoints[s - 1])
                    run_rerun((s - 1) % 2)
            run_adj(s % 2)
            if s > 0:
                main.wait_stream(side)

    def _capture_overlapped(self):
        """Graphs of the re-run into each buffer set and of the adjoint
        steps from each."""
        torch = self.torch
        self._rerun_g, self._adj_g = [], []
        for b in range(2):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._rerun_segment(b, self.u_side)
            self._rerun_g.append(g)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._adjoint_segment(b)
            self._adj_g.append(g)

    def _capture(self, u0):
        torch = self.torch
        self.u.copy_(u0)
        g_fwd, g_rev = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):            # warm-up outside capture
            self._forward_segment()
            self._reverse_segment()
        torch.cuda.current_stream().wait_stream(side)
        with torch.cuda.graph(g_fwd):
            self._forward_segment()
        self._fwd = g_fwd
        if self.overlap:
            self._capture_overlapped()
        else:
            with torch.cuda.graph(g_rev):
                self._reverse_segment()
            self._rev = g_rev

    def _sweep(self, u0, n_steps, run_fwd, run_rev, ckpt=None, run_rerun=None, run_adj=None):
        """The sweep's work, in the order a single run issues it: zeroed
        accumulators, the forward segments keeping each start state, then
        the reverse segments last first.  ckpt (n_seg, 5, nIC): checkpoint
        buffer (else fresh tensors)."""
        u = self.u
        u.copy_(u0)
        self.J.zero_()
        self.ua.zero_()
        if self.bc_grad is not None:
            self.bc_grad.zero_()
        nseg = n_steps // self.interval
        checkpoints = []
        for s in range(nseg):
            if ckpt is not None:
                ckpt[s].copy_(u)
                checkpoints.append(ckpt[s])
            else:
                checkpoints.append(u.clone())
            run_fwd()
        if self.keep_final:
            if ckpt is not None:
                self.u_final.copy_(u)
            else:
                self.u_final = u.clone()
        if run_rerun is not None:
            if ckpt is None:
                for t in checkpoints:                # read on the second stream too
                    t.record_stream(self._side)
            self._reverse_overlapped(checkpoints, run_rerun, run_adj)
            return
        while checkpoints:
            u.copy_(checkpoints.pop())
            run_rev()

    def _run_whole(self, u0, n_steps):
        torch = self.torch
        if self._whole is None or self._whole[0] != n_steps:
            self._whole = None
            nseg = n_steps // self.interval
            ckpt = torch.empty((nseg,) + tuple(self.u.shape), dtype=self.u.dtype, device='cuda')
            uin = torch.empty_like(self.u)
            if self.keep_final:
                self.u_final = torch.empty_like(self.u)
            uin.copy_(u0)
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):        # warm-up outside capture
                self._forward_segment()
                self._reverse_segment()
            torch.cuda.current_stream().wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                if self.overlap:
                    self._sweep(uin, n_steps, self._forward_segment, self._reverse_segment, ckpt,
                                lambda b: self._rerun_segment(b, self.u_side), self._adjoint_segment)
                else:
                    self._sweep(uin, n_steps, self._forward_segment, self._reverse_segment, ckpt)
            self._whole = (n_steps, g, ckpt, uin)
        _, g, ckpt, uin = self._whole
        uin.copy_(u0)
        g.replay()
        torch.cuda.synchronize()
        return np.asarray(self.J.cpu().numpy()), self.ua

    def run(self, u0, n_steps):
        """Forward n_steps steps from the device state u0 (unchanged), then
        the reverse sweep.  Returns (J, ua): J the FP64 sum over steps of the
        objective (one entry per cell segment of the context, on the host),
        ua = d(obja * J)/du0 on the device (a buffer reused by the next
        run)."""
        if n_steps % self.interval:
            raise ValueError('n_steps must be a multiple of interval')
        # a graph replays the launches it recorded: captured graphs are
        # dropped when the context's objective state (the stage whose state
        # the objective reads) has changed since
        state = self.ctx.objective_state
        if self._captured_for != state:
            self._fwd = self._rev = self._whole = None
            self._rerun_g = self._adj_g = None
            self._captured_for = state
        if self.whole:
            if not self.graphs:
                raise ValueError('whole=True captures the sweep as one graph: it needs graphs=True')
            return self._run_whole(u0, n_steps)
        if self.graphs and self._fwd is None:
            self._capture(u0)
        run_fwd = self._fwd.replay if self.graphs else self._forward_segment
        if self.overlap:
            if self.graphs:
                run_rerun = lambda b: self._rerun_g[b].replay()
                run_adj = lambda b: self._adj_g[b].replay()
            else:
                run_rerun = lambda b: self._rerun_segment(b, self.u_side)
                run_adj = self._adjoint_segment
            self._sweep(u0, n_steps, run_fwd, None, None, run_rerun, run_adj)
        else:
            run_rev = self._rev.replay if self.graphs else self._reverse_segment
            self._sweep(u0, n_steps, run_fwd, run_rev)
        self.torch.cuda.synchronize()
        return np.asarray(self.J.cpu().numpy()), self.ua


def graph_sweep(ctx, u0, dt, n_steps, interval, obja=1.0, graphs=True, tapes=True, source=None, whole=False,
                overlap=None):
    """One sweep (GraphSweep(...).run(u0, n_steps))."""
    return GraphSweep(ctx, dt, interval, obja, graphs, tapes, source, whole=whole, overlap=overlap).run(u0, n_steps)
