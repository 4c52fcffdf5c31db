"""Stepping machinery shared by the 2D and 3D solvers: graph recordings of step blocks and
the advance loop around them.

Once a solver's dt inputs are chained (every launch of a step reads what the previous step
wrote to device memory), a block of steps is a fixed sequence of HIP launches on torch's current
stream. It is recorded once per *variant* (the things that change what the block launches:
its length, whether its last step writes a diagnostics row or folds a running-mean sample,
whether means are on, the smoke buffer parity) and replayed with no host work.

The advance loop is planned against the two device rings every step writes into: the
coefficient history (one row per step, ``history_capacity`` rows) and, when diagnostics are on,
the diagnostics ring (one row per ``block`` steps, on the step reaching each multiple of
``block``). Rows are copied to the host only when a ring is full and at the end of the call.
"""
from __future__ import annotations

import time
import warnings

import torch


class StepGraphs:
    """Recorded step blocks of one solver, keyed by variant, in one memory pool. The first
    recording or replay error switches graphs off for the solver's lifetime (``error`` keeps
    the message) and the solver runs the same launches without them."""

    def __init__(self, device: torch.device, enabled: bool):
        self.device = device
        self.on = bool(enabled)
        self.error: str | None = None
        self._recorded: dict = {}
        self._pool = None

    def __len__(self) -> int:
        return len(self._recorded)

    def forget(self, where=None) -> None:
        """Drop every recording, or only those whose variant satisfies ``where(variant)``."""
        if where is None:
            self._recorded.clear()
        else:
            self._recorded = {v: g for v, g in self._recorded.items() if not where(v)}

    def play(self, variant, launches, times: int = 1, after_record=None, after_replay=None) -> bool:
        """Replay ``variant`` ``times`` times, first recording ``launches()`` if it has no
        recording yet. A recording executes nothing, so host bookkeeping that ``launches()``
        advanced is put back by ``after_record()``; ``after_replay()`` runs after each replay.
        Returns False (graphs now off, nothing of this call executed beyond completed replays)
        when recording or replay raises."""
        if not self.on:
            return False
        try:
            graph = self._recorded.get(variant)
            if graph is None:
                graph = self._record(launches)
                self._recorded[variant] = graph
                if after_record is not None:
                    after_record()
            for _ in range(times):
                graph.replay()
                if after_replay is not None:
                    after_replay()
        except Exception as exc:  # noqa: BLE001 - any HIP graph failure: the launches still run eagerly
            self._give_up(exc)
            return False
        return True

    def _record(self, launches) -> torch.cuda.CUDAGraph:
        if self._pool is None:
            self._pool = torch.cuda.graph_pool_handle()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, pool=self._pool):
            launches()
        return graph

    def _give_up(self, exc: Exception) -> None:
        self.error = f"{type(exc).__name__}: {exc}"
        self.on = False
        self._recorded.clear()
        warnings.warn(f"HIP graph path disabled, stepping without graphs ({self.error})",
                      RuntimeWarning, stacklevel=5)
        torch.cuda.synchronize(self.device)


class BlockStepping:
    """The ``advance`` loop of a solver. The solver class sets ``block``, ``history_capacity``
    and ``diag_capacity`` and provides

    - ``_graphs`` (:class:`StepGraphs`), ``_fresh`` (dt inputs chained), ``_diag`` (the device
      diagnostics ring or None), the ring counters ``_rows``/``_flushed`` and
      ``_diag_rows``/``_diag_flushed``, ``_diag_steps`` and ``steps``;
    - ``_blocks_allowed()``: whether whole recorded blocks may start at the current step;
    - ``_play_blocks(times)``: replay ``times`` blocks of ``block`` steps (False: not done);
    - ``_play_one(diag)`` / ``_launch_one(diag)``: one step, recorded or eager, writing a
      diagnostics row when ``diag``;
    - ``_flush()``: copy unread ring rows to the host.

    Each call appends ``(first step, last step, wall seconds)`` to ``_clock``."""

    block = 10
    history_capacity = 4096
    diag_capacity = 1024

    def _free_rows(self) -> int:
        """Steps that fit before a device ring would overwrite unread rows."""
        free = self.history_capacity - (self._rows - self._flushed)
        if self._diag is not None:
            free = min(free, self.block * (self.diag_capacity - (self._diag_rows - self._diag_flushed)))
        return free

    def _counted(self, steps: int, diag_rows_at=()) -> None:
        """Advance the host step and ring counters over ``steps`` completed steps, with a
        diagnostics row written on each step number in ``diag_rows_at``."""
        self.steps += steps
        self._rows += steps
        for s in diag_rows_at:
            self._diag_steps.append(s)
            self._diag_rows += 1

    def _single(self) -> None:
        """One step; the first after an edit runs eagerly (it forms its own dt inputs)."""
        diag = self._diag is not None and (self.steps + 1) % self.block == 0
        if not (self._fresh and self._play_one(diag)):
            self._launch_one(diag)
        self._counted(1, (self.steps + 1,) if diag else ())

    def advance(self, n: int) -> None:
        """Advance ``n`` steps with no autograd and no host synchronisation between flushes:
        recorded blocks where they may start, single steps elsewhere. A diagnostics row lands
        on the step reaching each multiple of ``block`` on every path, so recorded and eager
        runs produce the same rows."""
        todo = int(n)
        if todo <= 0:
            return
        first, clock = self.steps, time.perf_counter()
        with torch.no_grad():
            while todo > 0:
                free = self._free_rows()
                blocks = min(todo, free) // self.block if self._fresh and self._blocks_allowed() else 0
                if blocks and self._play_blocks(blocks):
                    span = blocks * self.block
                    marks = range(self.steps + self.block, self.steps + span + 1, self.block)
                    self._counted(span, marks if self._diag is not None else ())
                    todo -= span
                elif free > 0:
                    self._single()
                    todo -= 1
                if self._free_rows() <= 0:
                    self._flush()
            self._flush()
        self._clock.append((first, self.steps, time.perf_counter() - clock))
        if len(self._clock) > 10000:
            del self._clock[:5000]
