"""The verifier's file format and method choice, on the CPU: the parse
rules of cov/verify_cov.py:60-92 as cov/test_verify.py:125-141 checks them
(comments, blank lines, repeated words, separated integers, base-36
digits), the rejected inputs, and the ball volume."""
import math

import pytest

from coldcase_amd import verify_cov as V


def _write(tmp_path, text):
    p = tmp_path / 'c.txt'
    p.write_text(text)
    return str(p)


def test_comments_blank_lines_and_repeats(tmp_path):
    code, rep = V.parse_code(_write(tmp_path, '# comment\n0000\n0111\n\n0222\n0111\n'), 3, 4)
    assert code == [(0, 0, 0, 0), (0, 1, 1, 1), (0, 2, 2, 2)] and rep == 1


def test_separated_and_base36(tmp_path):
    code, _ = V.parse_code(_write(tmp_path, '0 1 2 2\n1,1,1,1\n2\t0\t1\t1  # trailing comment\n'), 3, 4)
    assert code == [(0, 1, 2, 2), (1, 1, 1, 1), (2, 0, 1, 1)]
    code, _ = V.parse_code(_write(tmp_path, '0a9\n0A8\n'), 12, 3)
    assert code == [(0, 10, 9), (0, 10, 8)]
    code, _ = V.parse_code(_write(tmp_path, '37 0 1\n'), 40, 3)
    assert code == [(37, 0, 1)] and V.fmt_word(code[0]) == '37 0 1' and V.fmt_word((0, 10, 9)) == '0a9'


@pytest.mark.parametrize('text', ['012\n', '0123\n0a23\n', '0 1 x 2\n', '0 1 2 3 4\n', '# only a comment\n\n',
                                  '0124\n'])
def test_rejected_files(tmp_path, text):
    with pytest.raises(SystemExit):
        V.parse_code(_write(tmp_path, text), 4, 4)


def test_ball_volume_and_method_choice():
    assert V.ball_volume(5, 7, 2) == 1 + 7 * 4 + 21 * 16 == 365
    assert V.ball_volume(6, 9, 4) == 90196
    assert V.ball_volume(3, 4, 9) == 3 ** 4
    assert sum(math.comb(7, i) * 6 ** i for i in range(5)) == V.ball_volume(7, 7, 4)
    assert V.choose_method(7, 5, 2, 400, 'auto') == 'pure'
    assert V.choose_method(10, 6, 4, 2784, 'auto') == 'numpy'
    assert V.choose_method(7, 5, 2, 400, 'numpy') == 'numpy'
