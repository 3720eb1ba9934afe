"""Write tests/golden/spx.npz: the S&P 500 series the reference's spx model
is built on (pypomp/data/SPX.csv, a data file of the reference), reduced to
what the model reads, as the reference's loader reduces it (spx.py:23-34):
time = days since the first date, y = log(Close / previous Close), the first
row (no previous close) dropped.  Inputs to the filters, not expected outputs.

Usage: python tests/golden/make_spx_fixture.py [--reference /root/reference]
"""

import argparse
from pathlib import Path

import numpy as np
import pandas as pd

HERE = Path(__file__).resolve().parent


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    raw = pd.read_csv(Path(args.reference) / "pypomp" / "data" / "SPX.csv")
    date = pd.to_datetime(raw["Date"])
    time = (date - date.min()).dt.days.astype(float).to_numpy()
    close = raw["Close"].to_numpy(dtype=np.float64)
    y = np.log(close[1:] / close[:-1])
    np.savez_compressed(HERE / "spx.npz", time=time[1:], y=y)
    print(f"spx.npz: {y.shape[0]} returns, times {time[1]}..{time[-1]}")


if __name__ == "__main__":
    main()
