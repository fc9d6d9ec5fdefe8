"""Adam reconstruction loop of 3DGSMR (trainers/trainer.py:237-283, :294-439)
around any voxelizer with the TileVoxelizer call."""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import torch

from .data import MRIProblem, fft3c, ifft3c
from .metrics import nmse, psnr
from .model import GaussianSet


@dataclass
class ReconConfig:
    lr_position: float = 1e-3
    lr_scale: float = 5e-3
    lr_rotation: float = 1e-3
    lr_density: float = 1e-2
    gamma: float = 0.999            # ExponentialLR per step
    max_grad_norm: float = 1.0      # 0 disables clipping
    densify: bool = False
    densify_from: int = 100
    densify_every: int = 100
    densify_until: int = 2500
    grad_threshold: float = 0.01
    scale_threshold: float = 0.01
    opacity_threshold: float = 0.01
    max_scale: float = 0.5
    max_points: int = 400_000
    clone: bool = True
    hard_dc: bool = True            # evaluate the data-consistent volume (trainers/trainer.py:237-243)

    @classmethod
    def single_lr(cls, lr, **kw):
        return cls(lr_position=lr, lr_scale=lr, lr_rotation=lr, lr_density=lr, **kw)


@dataclass
class History:
    rows: list = field(default_factory=list)

    def time_to(self, key, threshold):
        for r in self.rows:
            if r.get(key) is not None and r[key] <= threshold:
                return r["train_s"]
        return None


class Reconstructor:
    def __init__(self, problem: MRIProblem, gaussians: GaussianSet, voxelizer, config: ReconConfig):
        self.p = problem
        self.g = gaussians
        self.vox = voxelizer
        self.cfg = config
        self.step_count = 0
        self._pos_grad = None
        self._build_optimizer(None)

    # ----------------------------------------------------------- optimizer --
    def _groups(self):
        c = self.cfg
        return [{"params": [self.g.positions], "lr": c.lr_position},
                {"params": [self.g.log_scales], "lr": c.lr_scale},
                {"params": [self.g.rotations], "lr": c.lr_rotation},
                {"params": [self.g.density_real], "lr": c.lr_density},
                {"params": [self.g.density_imag], "lr": c.lr_density}]

    def _build_optimizer(self, row_map, old=None):
        lrs = None if old is None else [grp["lr"] for grp in old.param_groups]
        self.opt = torch.optim.Adam(self._groups())
        if lrs is not None:
            for grp, lr in zip(self.opt.param_groups, lrs):
                grp["lr"] = lr
                grp["initial_lr"] = lr
        if old is not None and row_map is not None:
            # Transplant Adam moments of surviving rows, zeros for new rows
            # (trainers/trainer.py:354-389).
            valid = row_map >= 0
            src = row_map.clamp_min(0)
            for og, ng in zip(old.param_groups, self.opt.param_groups):
                op, np_ = og["params"][0], ng["params"][0]
                st = old.state.get(op, {})
                if "exp_avg" not in st:
                    continue
                new_st = {"step": st["step"].clone() if torch.is_tensor(st["step"]) else st["step"]}
                for key in ("exp_avg", "exp_avg_sq"):
                    buf = st[key][src].clone()
                    buf[~valid] = 0
                    new_st[key] = buf
                self.opt.state[np_] = new_st
        self.sched = torch.optim.lr_scheduler.ExponentialLR(self.opt, gamma=self.cfg.gamma)

    # ------------------------------------------------------------ forward --
    def volume(self) -> torch.Tensor:
        out = self.vox(self.g.positions, self.g.scales(), self.g.rotations, self.g.density())
        return torch.complex(out[0], out[1]) if isinstance(out, tuple) else out

    def loss(self, vol: torch.Tensor) -> torch.Tensor:
        pred = fft3c(self.p.coils * vol.unsqueeze(0))
        diff = pred * self.p.mask.unsqueeze(0) - self.p.kspace
        return diff.real.square().sum() + diff.imag.square().sum()

    def hard_dc(self, vol: torch.Tensor) -> torch.Tensor:
        k = fft3c(self.p.coils * vol.unsqueeze(0))
        m = self.p.mask.unsqueeze(0)
        return torch.sum(torch.conj(self.p.coils) * ifft3c(k * (1.0 - m) + self.p.kspace), dim=0)

    def step(self) -> float:
        self.opt.zero_grad(set_to_none=True)
        loss = self.loss(self.volume())
        loss.backward()
        self._pos_grad = None if self.g.positions.grad is None else self.g.positions.grad.detach()
        if self.cfg.max_grad_norm > 0:
            torch.nn.utils.clip_grad_norm_(self.g.param_list(), self.cfg.max_grad_norm)
        self.opt.step()
        self.sched.step()
        self.step_count += 1
        if self.cfg.densify:
            self._density_control()
        return loss.detach()

    def _density_control(self):
        c, it = self.cfg, self.step_count
        if it < c.densify_from or it > c.densify_until or it % c.densify_every != 0 or self._pos_grad is None:
            return
        if self.g.num_points >= c.max_points:
            return
        gnorm = self._pos_grad.norm(dim=-1)
        scales = self.g.scales()
        n_split = int(((gnorm > c.grad_threshold) & (scales.max(-1)[0] > c.scale_threshold)).sum())
        if n_split and self.g.num_points + n_split <= c.max_points:
            old = self.opt
            cnt, row_map = self.g.densify_and_split(gnorm, c.grad_threshold, c.scale_threshold)
            self._build_optimizer(row_map, old)
        elif c.clone:
            n_clone = int(((gnorm > c.grad_threshold) & (scales.max(-1)[0] <= c.scale_threshold)).sum())
            if n_clone and self.g.num_points + n_clone <= c.max_points:
                old = self.opt
                cnt, row_map = self.g.densify_and_clone(gnorm, c.grad_threshold, c.scale_threshold)
                self._build_optimizer(row_map, old)
        with torch.no_grad():
            keep = ~((self.g.density().abs() < c.opacity_threshold) | (self.g.scales().max(-1)[0] > c.max_scale))
        if int(keep.sum()) >= 100 and int((~keep).sum()) > 0:
            old = self.opt
            _, row_map = self.g.prune(keep)
            self._build_optimizer(row_map, old)
        self._pos_grad = None

    @torch.no_grad()
    def evaluate(self):
        vol = self.volume()
        if self.cfg.hard_dc:
            vol = self.hard_dc(vol)
        return {"nmse": nmse(vol, self.p.target), "psnr": psnr(vol, self.p.target)}

    def run(self, steps: int, eval_every: int = 50, target_nmse: float | None = None, log=None) -> History:
        """Trains for `steps` Adam steps; evaluates every `eval_every` steps (the
        evaluation time is excluded from the training clock).  Stops early when
        `target_nmse` is reached."""
        hist = History()
        sync = torch.cuda.synchronize if self.p.target.is_cuda else (lambda: None)
        train_s = 0.0
        m = self.evaluate()
        hist.rows.append({"step": 0, "train_s": 0.0, "loss": None, "n": self.g.num_points, **m})
        for i in range(1, steps + 1):
            sync()
            t0 = time.perf_counter()
            loss = self.step()
            sync()
            train_s += time.perf_counter() - t0
            if i % eval_every == 0 or i == steps:
                m = self.evaluate()
                row = {"step": i, "train_s": train_s, "loss": float(loss), "n": self.g.num_points, **m}
                hist.rows.append(row)
                if log:
                    log(row)
                if target_nmse is not None and m["nmse"] <= target_nmse:
                    break
        return hist
